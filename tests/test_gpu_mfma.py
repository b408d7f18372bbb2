"""BF16 MFMA direct forms (plain and split-BF16) against the FP64 oracle.

Bars: split-BF16 (BF16X3) is an FP32-level path: potentials within 2e-6 and
one step within 1e-5 of FP64, like the FP32 forms.  Plain BF16 rounds every
state value and kernel weight to 8 bits (relative 2^-9): its error is
reported and held to a sanity bound (potential 2e-3, step 2e-2), not to the
FP32 bar.  Both forms are bit-deterministic.
"""
import numpy as np
import pytest

import oracle as O
from lenia_cases import (BF16, BF16X3, DIRECT, FFT, INIT_NONE, INIT_PATCH, config2, exp_shell,
                         gauss_shell, rand_state, ring)

pytestmark = pytest.mark.gpu

POT_TOL = {BF16X3: 2e-6, BF16: 2e-3}
STEP_TOL = {BF16X3: 1e-5, BF16: 2e-2}
NAMES = {BF16: "bf16", BF16X3: "bf16x3", DIRECT: "fp32-direct", FFT: "fp32-fft"}


@pytest.fixture(scope="module")
def ctx():
    import lenia

    c = lenia.Context(0)
    yield c
    c.close()


def err(a, b):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return float(np.abs(d).max()), float(np.sqrt(np.mean(d * d)))


def world(W, H, kernels, conv, **kw):
    d = dict(width=W, height=H, convMode=conv, growthMode=0, dt=0.1, radius=13, kernels=kernels,
             initMode=INIT_NONE)
    d.update(kw)
    return d


CASES = [
    ("shell_R5", 64, 32, [gauss_shell(5)]),
    ("exp_R13_ragged", 200, 72, [exp_shell(13)]),
    ("ring3_R7", 96, 60, [ring(6, [1], .15, .015, .5), ring(7, [1, .5], .26, .036, .3),
                          ring(5.5, [.5, 1, .25], .12, .012, .2)]),
    ("exp_R32", 160, 100, [exp_shell(32)]),
    # width not a multiple of 4: the state tile is staged one column per load
    ("exp_R9_width198", 198, 70, [exp_shell(9)]),
]


@pytest.mark.parametrize("conv", [BF16X3, BF16], ids=["bf16x3", "bf16"])
@pytest.mark.parametrize("name,W,H,ks", CASES, ids=[c[0] for c in CASES])
def test_potential_and_step(ctx, conv, name, W, H, ks):
    import lenia

    d = world(W, H, ks, conv)
    A = rand_state(W, H, seed=11)
    with lenia.World(ctx, d) as w:
        w.write(A)
        for k, kd in enumerate(ks):
            u = w.potential(k)
            want = O.potential_direct(A.astype(np.float64), O.kernel_window(kd))
            mx, rms = err(u, want)
            print(f"{name} {NAMES[conv]} k={k}: potential max-abs {mx:.2e} rms {rms:.2e}")
            assert mx <= POT_TOL[conv]
        w.step(1)
        got = w.read()
    want = O.step(d, A.astype(np.float64), use_fft=False)
    mx, rms = err(got, want)
    print(f"{name} {NAMES[conv]}: step max-abs {mx:.2e} rms {rms:.2e}")
    assert mx <= STEP_TOL[conv]


@pytest.mark.parametrize("conv", [BF16X3, BF16], ids=["bf16x3", "bf16"])
def test_deterministic_and_frame_semantics(ctx, conv):
    import lenia

    d = dict(config2(conv=conv), width=320, height=180)
    outs = []
    for _ in range(2):
        with lenia.World(ctx, d) as w:
            w.reset(4242)
            w.frame(0, 3)
            w.step(5)
            outs.append(w.read().tobytes())
    assert outs[0] == outs[1]


def test_config2_error_report(ctx):
    """1280x720, one exp-bump kernel R=32, full-frame seeded noise: max-abs and
    RMS against the FP64 oracle after 1, 10 and 100 steps for every form, and
    the per-step error from each form's own state at step 99"""
    import lenia

    base = config2()
    m = O.module_seed(7777)
    A0 = O.init_patch(1280, 720, m, 0, 0, 1280, 720)
    ref = {0: A0}
    A = A0
    for t in range(1, 101):
        A = O.step(base, A, use_fft=True)
        if t in (1, 10, 99, 100):
            ref[t] = A
    for conv in (BF16, BF16X3, DIRECT, FFT):
        d = dict(base, convMode=conv)
        with lenia.World(ctx, d) as w:
            w.reset(7777)
            w.frame(0, 1)
            res = {1: w.read()}
            w.step(9)
            res[10] = w.read()
            w.step(89)
            s99 = w.read()
            w.step(1)
            res[100] = w.read()
        line = []
        for t in (1, 10, 100):
            mx, rms = err(res[t], ref[t])
            line.append(f"step {t}: max {mx:.2e} rms {rms:.2e}")
        own = O.step(base, s99.astype(np.float64), use_fft=True)
        mx1, rms1 = err(res[100], own)
        print(f"config2 {NAMES[conv]}: " + "; ".join(line) +
              f"; per-step at 100: max {mx1:.2e} rms {rms1:.2e}")
        if conv != BF16:
            assert mx1 <= 1e-5
            mx, _ = err(res[1], ref[1])
            assert mx <= 1e-5
        else:
            assert mx1 <= 2e-2


def test_window_limits(ctx):
    """the state tile of a workgroup must fit 64 KB of LDS: split-BF16 serves
    h <= 32, plain BF16 h <= 41; larger windows are rejected (the FP32 forms
    serve them)"""
    import lenia

    with pytest.raises(lenia.LeniaError) as ei:
        lenia.World(ctx, world(256, 256, [exp_shell(33)], BF16X3))
    assert ei.value.code == -2
    with pytest.raises(lenia.LeniaError) as ei:
        lenia.World(ctx, world(256, 256, [exp_shell(42)], BF16))
    assert ei.value.code == -2
    A = rand_state(256, 256, seed=3)
    for conv, R in ((BF16X3, 32), (BF16, 41)):
        kd = exp_shell(R)
        with lenia.World(ctx, world(256, 256, [kd], conv)) as w:
            w.write(A)
            u = w.potential(0)
        want = O.potential_direct(A.astype(np.float64), O.kernel_window(kd))
        mx, _ = err(u, want)
        print(f"{NAMES[conv]} R={R}: potential max-abs {mx:.2e}")
        assert mx <= POT_TOL[conv]
