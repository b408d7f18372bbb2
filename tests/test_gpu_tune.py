"""Launch-shape autotuning (lenia_world_autotune / lenia_batch_autotune):
tuning changes only which threads run a butterfly, never the arithmetic, so a
tuned world or batch gives the same bytes as an untuned one, and the tuning
runs leave the state and step clock where they were."""
import numpy as np
import pytest

from lenia_cases import config1, config4, sweep_worlds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import lenia

    c = lenia.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("desc", [config4(), config1(W=512, H=256)], ids=["256x256_K2", "512x256_K3"])
def test_world_autotune_keeps_bits_and_state(ctx, desc):
    import lenia

    with lenia.World(ctx, desc) as plain, lenia.World(ctx, desc) as tuned:
        for w in (plain, tuned):
            w.reset(2024)
            w.frame(0, 1)
            w.step(5)
        before = tuned.read()
        ms = tuned.autotune(8)
        after = tuned.read()
        assert np.array_equal(before, after)
        print(f"{desc['width']}x{desc['height']}: tuned {tuned.resolved()} {ms * 1e3:.1f} us/step, "
              f"default {plain.resolved()}")
        assert ms > 0
        for w in (plain, tuned):
            w.step(40)
        assert plain.read().tobytes() == tuned.read().tobytes()
        # the frame clock continued where it was: frame semantics still match
        plain.frame(100, 2)
        tuned.frame(100, 2)
        assert plain.read().tobytes() == tuned.read().tobytes()


def test_batch_autotune_keeps_bits(ctx):
    import lenia

    descs = sweep_worlds(16, seed=11)
    seeds = [500 + i for i in range(len(descs))]
    worlds = [lenia.World(ctx, d) for d in descs]
    alone = [lenia.World(ctx, d) for d in descs]
    try:
        for ws in (worlds, alone):
            for w, s in zip(ws, seeds):
                w.reset(s)
                w.frame(0, 1)
        with lenia.Batch(ctx, worlds) as b:
            ms = b.autotune(6)
            print(f"batch of {len(descs)} tuned: {ms * 1e3:.1f} us/step")
            b.step(16, 8)
        for w in alone:
            w.step(16)
        for i, (w, a) in enumerate(zip(worlds, alone)):
            assert w.read().tobytes() == a.read().tobytes(), i
    finally:
        for w in worlds + alone:
            w.close()


def test_row_pass_workgroup_sizes_same_bits(ctx):
    """the headline world stepped with each row-pass workgroup size the
    library offers (1024 threads: side-by-side kernel transforms only) gives
    the same bytes"""
    import lenia

    out = {}
    for rt in (128, 256, 512, 1024):
        d = dict(config1(), rowThreads=rt, rowPairsPerTile=1)
        with lenia.World(ctx, d) as w:
            assert w.resolved()["rowThreads"] == rt
            w.reset(7777)
            w.frame(0, 1)
            w.step(12)
            out[rt] = w.read().tobytes()
    assert len(set(out.values())) == 1


@pytest.mark.gpu
def test_row_pass_1024_register_cap(ctx):
    """1024-thread row passes hold at most 12 cells per thread, and tiles of at
    most 4 run a variant sized for them.  A 2048-wide world with four kernels
    and two row pairs per tile (4 x 2 x 2048 > 12 x 1024, within the LDS) is
    refused with an error, not run with cells left out; one row pair per tile
    (the 12-cell variant) and a 4096-wide world with three kernels (12 x 1024
    exactly) give the same bytes as 512-thread workgroups"""
    import lenia
    from lenia_cases import ring

    k4 = [dict(k) for k in config1()["kernels"]] + [ring(24, [1.0], 0.2, 0.02, 0.1)]
    wide4 = dict(config1(2048, 128), kernels=k4, initRect=(0, 0, 0, 0))
    with pytest.raises(lenia.LeniaError, match="row-pass shape"):
        lenia.World(ctx, dict(wide4, rowThreads=1024, rowPairsPerTile=2))
    for d in (wide4, dict(config1(4096, 64), initRect=(0, 0, 0, 0))):
        out = {}
        for rt in (512, 1024):
            with lenia.World(ctx, dict(d, rowThreads=rt, rowPairsPerTile=1)) as w:
                assert w.resolved()["rowThreads"] == rt
                w.reset(7777)
                w.frame(0, 1)
                w.step(6)
                out[rt] = w.read().tobytes()
        assert out[512] == out[1024], d["width"]


@pytest.mark.gpu
def test_column_pass_lean_variant_same_bits(ctx):
    """column tiles of at most 8 values per thread run a variant sized for
    them (3 x 512 cells at 256 threads); 64-thread tiles of the same world run
    the full-size variant: both give the same bytes"""
    import lenia

    out = {}
    for ct in (64, 256):
        d = dict(config1(), colThreads=ct, colsPerTile=1)
        with lenia.World(ctx, d) as w:
            assert w.resolved()["colThreads"] == ct
            w.reset(7777)
            w.frame(0, 1)
            w.step(12)
            out[ct] = w.read().tobytes()
    assert out[64] == out[256]


# every radix sequence with a compile-time kernel (packed: 4 bits per stage,
# first stage lowest, F = 16); the default plans and the tuner's alternatives
COMPILED_PLANS = {
    64: [0x88, 0x4F, 0xF4], 128: [0x8F, 0xF8, 0x844], 256: [0xFF, 0x488, 0x884],
    512: [0x888, 0x48F, 0xF84], 1024: [0x88F, 0xF88, 0x4FF, 0xFF4],
    2048: [0x8FF, 0xFF8], 4096: [0xFFF, 0x8888], 96: [0x348, 0x384], 192: [0x388, 0x34F],
    384: [0x38F, 0x3F8, 0x3448],
}


@pytest.mark.parametrize("n", sorted(COMPILED_PLANS))
def test_compiled_plans_match_oracle(ctx, n):
    """each compiled radix plan of length n, as the row transform (width n) and
    as the column transform (height n), one step within 1e-5 of the FP64 oracle"""
    import lenia
    import oracle as O
    from lenia_cases import FFT, INIT_NONE, rand_state, ring

    ks = [ring(8, [1.0], 0.2, 0.03, 0.6), ring(6, [1.0, 0.5], 0.3, 0.05, 0.4)]
    worst = 0.0
    for plan in COMPILED_PLANS[n]:
        for W, H, key in ((n, 64, "fftPlanW"), (64, n, "fftPlanH")):
            d = dict(width=W, height=H, convMode=FFT, growthMode=0, dt=0.1, radius=8, kernels=ks,
                     initMode=INIT_NONE, **{key: plan})
            A = rand_state(W, H, seed=plan & 0xFFF)
            with lenia.World(ctx, d) as w:
                assert w.resolved()[key] == plan
                w.write(A)
                w.step(1)
                got = w.read()
            want = O.step(d, A.astype(np.float64), use_fft=True)
            err = float(np.abs(got - want).max())
            worst = max(worst, err)
            assert err <= 1e-5, (n, hex(plan), key, err)
    print(f"length {n}: {len(COMPILED_PLANS[n])} plans, worst one-step error {worst:.2e}")


def test_plan_autotune_shared_by_worlds_and_batches(ctx):
    """lenia_plan_autotune records the fastest compiled plans of a world's
    width and height (for the height, a plan with a wave kernel within 10% of
    the fastest); worlds created afterwards take them, a batch of such
    worlds equals the same worlds stepped alone bit for bit, and the step is
    within 1e-5 of the FP64 oracle.  An explicit plan still wins."""
    import lenia
    import oracle as O
    from lenia_cases import INIT_NONE, rand_state

    lenia.plan_cache_clear()
    try:
        desc = config1(W=512, H=256)
        pw, ph, ms = lenia.plan_autotune(ctx, desc, 6)
        print(f"512x256 K=3: tuned plans rows 0x{pw:x} cols 0x{ph:x}, {ms * 1e3:.1f} us per step")
        assert pw in COMPILED_PLANS[512] and ph in COMPILED_PLANS[256]
        assert lenia.plan_cache_get(lenia.ROWS, 512) == pw
        assert lenia.plan_cache_get(lenia.COLS, 256) == ph
        descs = [dict(desc, initMode=INIT_NONE), dict(desc, initMode=INIT_NONE, dt=0.05)]
        states = [rand_state(512, 256, seed=s) for s in (1, 2)]
        alone = []
        for d, A in zip(descs, states):
            with lenia.World(ctx, d) as w:
                assert (w.resolved()["fftPlanW"], w.resolved()["fftPlanH"]) == (pw, ph)
                w.write(A)
                w.step(9)
                alone.append(w.read())
        worlds = [lenia.World(ctx, d) for d in descs]
        try:
            for w, A in zip(worlds, states):
                w.write(A)
            with lenia.Batch(ctx, worlds) as b:
                b.step(9)
            for w, want in zip(worlds, alone):
                assert w.read().tobytes() == want.tobytes()
        finally:
            for w in worlds:
                w.close()
        with lenia.World(ctx, descs[0]) as w:
            w.write(states[0])
            w.step(1)
            err = float(np.abs(w.read() - O.step(descs[0], states[0].astype(np.float64),
                                                 use_fft=True)).max())
        print(f"tuned plans, one step vs FP64 oracle: max-abs {err:.2e}")
        assert err <= 1e-5
        other = [p for p in COMPILED_PLANS[512] if p != pw][0]
        with lenia.World(ctx, dict(descs[0], fftPlanW=other)) as w:
            assert w.resolved()["fftPlanW"] == other
    finally:
        lenia.plan_cache_clear()
