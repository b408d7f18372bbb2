"""GPU reverse pass of the differentiable rollout, through the C ABI.

The HIP gradient (FP32 arithmetic, FP64 parameter reductions) is checked
against the oracle's FP64 hand-written gradient, which tests/test_oracle.py
pins against torch autograd in float64 and central finite differences.
Tolerance: relative L2 error 2e-3 on the state gradient and 2e-3 relative
to the largest parameter gradient on each parameter (FP32 forward states
and FP32 spectral products against FP64; cells whose pre-clip value sits
within rounding of 0 or 1 can take the other side of the clip mask).
Bit-identity: checkpoint spacing and repeated calls do not change a bit.
"""
import numpy as np
import pytest

import oracle as O
from lenia_cases import DIRECT, FFT, INIT_NONE, INIT_PATCH, config4, ring

pytestmark = pytest.mark.gpu

GRAD_TOL = 2e-3


@pytest.fixture(scope="module")
def ctx():
    import lenia

    c = lenia.Context(0)
    yield c
    c.close()


def grad_world(conv=FFT, **kw):
    """24x20 torus, ring kernels R=4 (one shell) and R=5 (two shells)"""
    d = dict(width=24, height=20, convMode=conv, growthMode=0, dt=0.1, radius=4,
             initMode=INIT_NONE,
             kernels=[ring(4, [1.0], 0.2, 0.05, 0.6), ring(5, [1.0, 0.6], 0.25, 0.06, 0.4)])
    d.update(kw)
    return d


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def compare_grads(g, ref, K, label):
    e = rel_l2(g["A0"], ref["A0"])
    print(f"{label}: dA0 rel L2 {e:.3e}")
    assert e <= GRAD_TOL, (label, "A0", e)
    scale = max(np.abs(np.concatenate([ref["mu"], ref["sigma"], ref["weight"]])).max(),
                abs(ref["dt"]))
    for key in ("mu", "sigma", "weight"):
        d = np.abs(g[key] - ref[key]).max() / scale
        print(f"{label}: d{key} {g[key]} ref {ref[key]} rel {d:.3e}")
        assert d <= GRAD_TOL, (label, key, g[key], ref[key])
    d = abs(g["dt"] - ref["dt"]) / scale
    print(f"{label}: ddt {g['dt']} ref {ref['dt']} rel {d:.3e}")
    assert d <= GRAD_TOL, (label, "dt", g["dt"], ref["dt"])
    bscale = max(np.abs(ref["betas"]).max(), 1e-12)
    for k in range(K):
        d = np.abs(g["betas"][k] - ref["betas"][k]).max() / bscale
        print(f"{label}: dbetas[{k}] {g['betas'][k]} ref {ref['betas'][k]} rel {d:.3e}")
        assert d <= GRAD_TOL, (label, "betas", k)


def small_inputs(seed=9):
    rng = np.random.default_rng(seed)
    A0 = (0.2 + 0.6 * rng.random((20, 24))).astype(np.float32)
    gT = rng.standard_normal((20, 24)).astype(np.float32)
    return A0, gT


def test_rollout_forward_matches_oracle(ctx):
    import lenia

    w = lenia.World(ctx, grad_world())
    A0, _ = small_inputs()
    T = 6
    AT = w.rollout(A0, T)
    ref = O.rollout(grad_world(), A0.astype(np.float64), T)
    e = float(np.abs(AT - ref).max())
    print(f"rollout T={T} max-abs {e:.3e}")
    assert e <= T * 1e-5
    # the same as writing the state and stepping
    w.write(A0)
    w.step(T)
    assert np.array_equal(w.read(), AT)
    w.close()


@pytest.mark.parametrize("T,ckpt", [(1, 0), (6, 0), (6, 2), (7, 3)])
def test_gradient_matches_oracle(ctx, T, ckpt):
    import lenia

    w = lenia.World(ctx, grad_world())
    A0, gT = small_inputs()
    g = w.rollout_grad(A0, T, gT, ckpt_every=ckpt)
    ref = O.rollout_grad(grad_world(), A0.astype(np.float64), T, gT.astype(np.float64))
    compare_grads(g, ref, 2, f"24x20 T={T} ckpt={ckpt}")
    w.close()


def test_gradient_checkpoint_spacing_and_repeat_bit_identical(ctx):
    import lenia

    w = lenia.World(ctx, grad_world())
    A0, gT = small_inputs(3)
    T = 9
    outs = [w.rollout_grad(A0, T, gT, ckpt_every=c) for c in (1, 2, 3, 0, 9, 3)]
    for o in outs[1:]:
        assert np.array_equal(o["A0"], outs[0]["A0"])
        for key in ("mu", "sigma", "weight", "betas"):
            assert np.array_equal(o[key], outs[0][key]), key
        assert o["dt"] == outs[0]["dt"]
    w.close()


def test_gradient_reuses_forward_checkpoints(ctx):
    """rollout() keeps its checkpoints; rollout_grad(None, ...) reuses them and
    gives the same bits as a reverse pass that reruns the sweep from A0; a
    parameter change in between makes the reuse an error"""
    import lenia

    w = lenia.World(ctx, grad_world())
    A0, gT = small_inputs(6)
    T, c = 9, 3
    full = w.rollout_grad(A0, T, gT, ckpt_every=c)
    w.rollout(A0, T, ckpt_every=c)
    reuse = w.rollout_grad(None, T, gT, ckpt_every=c)
    assert np.array_equal(full["A0"], reuse["A0"])
    for key in ("mu", "sigma", "weight", "betas"):
        assert np.array_equal(full[key], reuse[key]), key
    assert full["dt"] == reuse["dt"]
    with pytest.raises(lenia.LeniaError) as ei:
        w.rollout_grad(None, T + 1, gT, ckpt_every=c)
    assert ei.value.code == -5
    w.rollout(A0, T, ckpt_every=c)
    w.set_params(grad_world(dt=0.12))
    with pytest.raises(lenia.LeniaError) as ei:
        w.rollout_grad(None, T, gT, ckpt_every=c)
    assert ei.value.code == -5
    w.close()


def seeded_state(ctx, desc, seed, steps=0):
    """the world's own seeded init (full-frame noise), then `steps` steps"""
    import lenia

    w = lenia.World(ctx, dict(desc, initMode=INIT_PATCH))
    w.reset(seed)
    w.frame(0)
    if steps:
        w.step(steps)
    A = w.read()
    w.close()
    return A


def test_gradient_config_256_two_kernels(ctx):
    """256x256, ring kernels R=16 (one and two shells), T=16 with checkpoints
    every 4, from one step after seeded full-frame noise"""
    import lenia

    desc = dict(config4(), initMode=INIT_NONE)
    A0 = seeded_state(ctx, desc, 1234)
    inside = float(((A0 > 0) & (A0 < 1)).mean())
    assert inside > 0.5, inside
    rng = np.random.default_rng(5)
    gT = rng.standard_normal(A0.shape).astype(np.float32)
    T = 16
    w = lenia.World(ctx, desc)
    g = w.rollout_grad(A0, T, gT, ckpt_every=4)
    ref = O.rollout_grad(desc, A0.astype(np.float64), T, gT.astype(np.float64))
    compare_grads(g, ref, 2, "256x256 T=16")
    w.close()


C4_GRAD_TOL = 1e-4  # observed 1.2e-6 (dA0 rel L2) and 4e-6 (parameters) on MI355X


def test_gradient_config4_T64_ckpt8(ctx):
    """The differentiable-rollout configuration at its stated size: 256x256,
    two ring kernels R=16, T=64 steps with checkpoints every 8 (recomputed
    segments), from one step after seeded full-frame noise, upstream gradient
    of an MSE-like loss.  Against the FP64 oracle gradient (which
    tests/test_oracle.py checks against FP64 central differences at this same
    size): dA0 relative L2 and every parameter gradient relative to the
    largest within 1e-4; the state at T is still alive (the gradient flows
    through the clip mask of live cells)."""
    import lenia

    desc = dict(config4(), initMode=INIT_NONE)
    A0 = seeded_state(ctx, desc, 1234)
    rng = np.random.default_rng(5)
    gT = rng.standard_normal(A0.shape).astype(np.float32)
    T, c = 64, 8
    with lenia.World(ctx, desc) as w:
        AT = w.rollout(A0, T, ckpt_every=c)
        g = w.rollout_grad(None, T, gT, ckpt_every=c)
    live = float(((AT > 0) & (AT < 1)).mean())
    print(f"256x256 T=64: {100 * live:.1f}% of cells live at T")
    assert live > 0.05
    ref = O.rollout_grad(desc, A0.astype(np.float64), T, gT.astype(np.float64))
    e = rel_l2(g["A0"], ref["A0"])
    print(f"256x256 T=64 ckpt=8: dA0 rel L2 {e:.3e}")
    assert e <= C4_GRAD_TOL
    scale = max(np.abs(np.concatenate([ref["mu"], ref["sigma"], ref["weight"]])).max(),
                abs(ref["dt"]), np.abs(ref["betas"]).max())
    for key in ("mu", "sigma", "weight", "betas"):
        d = float(np.abs(g[key] - ref[key]).max() / scale)
        print(f"256x256 T=64 ckpt=8: d{key} rel {d:.3e}  gpu {np.ravel(g[key])[:4]} "
              f"fp64 {np.ravel(ref[key])[:4]}")
        assert d <= C4_GRAD_TOL, key
    d = abs(g["dt"] - ref["dt"]) / scale
    print(f"256x256 T=64 ckpt=8: ddt rel {d:.3e}  gpu {g['dt']:.6f} fp64 {ref['dt']:.6f}")
    assert d <= C4_GRAD_TOL


def test_gradient_after_param_change(ctx):
    """set_params between calls: the reverse pass follows the new values"""
    import lenia

    w = lenia.World(ctx, grad_world())
    A0, gT = small_inputs(4)
    w.rollout_grad(A0, 4, gT)
    d2 = grad_world(dt=0.15)
    d2["kernels"][0]["mu"] = 0.21
    d2["kernels"][1]["betas"] = [0.8, 0.7]
    w.set_params(d2)
    g = w.rollout_grad(A0, 4, gT)
    ref = O.rollout_grad(d2, A0.astype(np.float64), 4, gT.astype(np.float64))
    compare_grads(g, ref, 2, "after set_params")
    w.close()


def test_gradient_unsupported_forms(ctx):
    import lenia

    A0, gT = small_inputs()
    for desc in (grad_world(conv=DIRECT), grad_world(noiseAmount=0.01),
                 grad_world(growthMode=1)):
        w = lenia.World(ctx, desc)
        with pytest.raises(lenia.LeniaError) as ei:
            w.rollout_grad(A0, 3, gT)
        assert ei.value.code == -2
        with pytest.raises(lenia.LeniaError):
            w.rollout(A0, 3)
        w.close()
    w = lenia.World(ctx, grad_world())
    with pytest.raises(lenia.LeniaError) as ei:
        w.rollout_grad(A0, 0, gT)
    assert ei.value.code == -1
    w.close()
