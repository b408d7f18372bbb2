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
