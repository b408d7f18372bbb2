"""The wave executor of the FFT-form passes against the workgroup-LDS executor.

Both executors run the same radix plan with the same butterflies, twiddles and
output placement, and the same gather, growth and spectrum split around them,
so every comparison here is byte for byte: single worlds of every length that
has a wave kernel (rows and columns), the frame-seeded noise / jitter / inject
terms, potentials, every waves-per-workgroup choice, batches mixing both
executors, and the rollout gradient, whose forward recomputation runs the
world's executor.  Parity of the numbers themselves with the FP64 oracle is
covered by the other GPU tests, which now run on the wave executor wherever a
plan has one (the library's default)."""
import numpy as np
import pytest

import lenia
from lenia_cases import FFT, INIT_PATCH, ring, sweep_worlds

pytestmark = pytest.mark.gpu

LDS, WAVE, SPLIT = 1, 2, 3


@pytest.fixture(scope="module")
def ctx():
    c = lenia.Context(0)
    yield c
    c.close()


def desc(W, H, K, mode=0, extras=False, **kw):
    R = min(12, (min(W, H) - 1) // 2)
    ks = [ring(R, [1.0, 0.5, 0.8][: 1 + k % 3], 0.15 + 0.03 * k, 0.02 + 0.005 * k, 1.0 / K)
          for k in range(K)]
    d = dict(width=W, height=H, convMode=FFT, growthMode=mode, dt=0.1, radius=R, kernels=ks,
             initMode=INIT_PATCH, initRect=(0, 0, 0, 0))
    if extras:
        d.update(noiseAmount=0.4, muJitter=0.3, injectAmount=0.8, injectCount=2)
    d.update(kw)
    return d


def run(ctx, d, steps=6, seed=4242, potentials=False):
    with lenia.World(ctx, d) as w:
        w.reset(seed)
        w.frame(3, 2)
        w.step(steps)
        out = [w.read()]
        if potentials:
            out += [w.potential(k) for k in range(len(d["kernels"]))]
        res = w.resolved()
    return out, res


# (W, H): every row length and every column length with a wave kernel
SIZES = [(64, 96), (96, 64), (128, 192), (192, 128), (256, 512), (512, 256), (1024, 128), (96, 1024)]


@pytest.mark.parametrize("W,H", SIZES)
@pytest.mark.parametrize("K", [1, 3])
def test_wave_equals_lds_executor(ctx, W, H, K):
    d = desc(W, H, K, mode=K % 2)
    a, ra = run(ctx, dict(d, rowEngine=LDS, colEngine=LDS), potentials=True)
    assert (ra["rowEngine"], ra["colEngine"]) == (LDS, LDS)
    for eng in (WAVE, SPLIT):
        b, rb = run(ctx, dict(d, rowEngine=eng, colEngine=eng), potentials=True)
        assert (rb["rowEngine"], rb["colEngine"]) == (eng, eng)
        assert (ra["fftPlanW"], ra["fftPlanH"]) == (rb["fftPlanW"], rb["fftPlanH"])
        for x, y in zip(a, b):
            assert x.tobytes() == y.tobytes(), eng
    live = int(np.count_nonzero((a[0] > 0) & (a[0] < 1)))
    print(f"{W}x{H} K={K}: wave and kernel-split == LDS executor after 6 steps ({live} live cells)")
    assert live > 0


# plans with a wave kernel that are not a length's default: (W, H, planW, planH)
EXPLICIT = [(384, 192, 0x3448, 0), (128, 384, 0, 0x3448)]


@pytest.mark.parametrize("W,H,pw,ph", EXPLICIT)
@pytest.mark.parametrize("K", [1, 3])
def test_wave_equals_lds_executor_explicit_plans(ctx, W, H, pw, ph, K):
    d = desc(W, H, K, mode=K % 2, fftPlanW=pw, fftPlanH=ph)
    a, ra = run(ctx, dict(d, rowEngine=LDS, colEngine=LDS), potentials=True)
    for eng in (WAVE, SPLIT):
        b, rb = run(ctx, dict(d, rowEngine=eng, colEngine=eng), potentials=True)
        assert (rb["rowEngine"], rb["colEngine"]) == (eng, eng)
        assert (ra["fftPlanW"], ra["fftPlanH"]) == (rb["fftPlanW"], rb["fftPlanH"])
        for x, y in zip(a, b):
            assert x.tobytes() == y.tobytes(), eng
    assert int(np.count_nonzero((a[0] > 0) & (a[0] < 1))) > 0


@pytest.mark.parametrize("W,H", [(512, 192), (96, 256)])
def test_wave_extras_and_mixed_executors(ctx, W, H):
    """noise / jitter / inject terms on the wave executor, and one pass on
    each executor"""
    d = desc(W, H, 2, extras=True)
    ref, _ = run(ctx, dict(d, rowEngine=LDS, colEngine=LDS), steps=9)
    for re_, ce in ((WAVE, WAVE), (WAVE, LDS), (LDS, WAVE), (SPLIT, SPLIT), (SPLIT, WAVE), (LDS, SPLIT)):
        got, _ = run(ctx, dict(d, rowEngine=re_, colEngine=ce), steps=9)
        assert got[0].tobytes() == ref[0].tobytes(), (re_, ce)


@pytest.mark.parametrize("threads", [64, 128, 256])
def test_wave_workgroup_sizes(ctx, threads):
    d = desc(256, 192, 2)
    ref, _ = run(ctx, dict(d, rowEngine=LDS, colEngine=LDS))
    got, r = run(ctx, dict(d, rowEngine=WAVE, colEngine=WAVE, rowThreads=threads, colThreads=threads))
    assert (r["rowThreads"], r["colThreads"]) == (threads, threads)
    assert got[0].tobytes() == ref[0].tobytes()


def test_wave_rejects_what_it_cannot_run(ctx):
    # 384 (16 8 3) has no wave kernel; an LDS tile shape with the wave executor
    with pytest.raises(lenia.LeniaError):
        lenia.World(ctx, desc(384, 64, 1, rowEngine=WAVE))
    with pytest.raises(lenia.LeniaError):
        lenia.World(ctx, desc(256, 64, 1, rowEngine=WAVE, rowPairsPerTile=2))
    with pytest.raises(lenia.LeniaError):
        lenia.World(ctx, desc(256, 64, 1, colEngine=WAVE, colThreads=512))
    with pytest.raises(lenia.LeniaError):
        lenia.World(ctx, desc(256, 64, 1, rowEngine=SPLIT, rowThreads=128))
    with pytest.raises(lenia.LeniaError):
        lenia.World(ctx, desc(256, 64, 1, rowEngine=4))
    # the library's choice: kernel-split wave columns where the plan has a
    # wave kernel, LDS rows
    with lenia.World(ctx, desc(384, 256, 1)) as w:
        r = w.resolved()
        assert (r["rowEngine"], r["colEngine"]) == (LDS, SPLIT)
    with lenia.World(ctx, desc(384, 384, 1)) as w:
        r = w.resolved()
        assert (r["rowEngine"], r["colEngine"]) == (LDS, LDS)


def test_batch_wave_groups_equal_lds_and_alone(ctx):
    """48 sweep worlds (all seven side lengths, 1-3 kernels): the batch on the
    wave executor (classes of 8, 16 and 24 values per lane plus the LDS group
    for length 384), on the LDS executor only, and each world alone"""
    descs = sweep_worlds(48)
    states = {}
    for waves in ((4, 4), (1, 2), (3, 3), (0, 0)):
        ws = [lenia.World(ctx, d) for d in descs]
        try:
            for i, w in enumerate(ws):
                w.reset(100 + i)
            with lenia.Batch(ctx, ws) as b:
                b.set_waves(*waves)
                assert b.waves() == waves
                b.step(12, graph_steps=4)
                states[waves] = [w.read() for w in ws]
        finally:
            for w in ws:
                w.close()
    alone = []
    for i, d in enumerate(descs):
        with lenia.World(ctx, d) as w:
            w.reset(100 + i)
            w.step(12)
            alone.append(w.read())
    for i in range(len(descs)):
        for waves in ((4, 4), (1, 2), (3, 3)):
            assert states[waves][i].tobytes() == states[(0, 0)][i].tobytes(), (i, waves)
        assert alone[i].tobytes() == states[(0, 0)][i].tobytes(), i


def test_rollout_gradient_same_on_both_executors(ctx):
    """the reverse pass recomputes each segment's forward steps on the world's
    executor; gradients are identical bit for bit"""
    d = desc(256, 128, 2)
    rng = np.random.default_rng(5)
    A0 = rng.random((128, 256)).astype(np.float32)
    gT = rng.standard_normal((128, 256)).astype(np.float32)
    out = []
    for eng in (LDS, WAVE, SPLIT):
        with lenia.World(ctx, dict(d, rowEngine=eng, colEngine=eng)) as w:
            g = w.rollout_grad(A0, 12, gT, ckpt_every=4)
        out.append(g)
    for g in out[1:]:
        for key in out[0]:
            assert np.asarray(out[0][key]).tobytes() == np.asarray(g[key]).tobytes(), key
