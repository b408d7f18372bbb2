"""Batched stepping: every world of a batch gives the same bytes as the same
world stepped alone, for the FFT and direct forms, with and without graph
replay and with the per-frame noise / jitter / inject terms."""
import numpy as np
import pytest

from lenia_cases import DIRECT, FFT, INIT_NONE, exp_shell, rand_state, ring, sweep_worlds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import lenia

    c = lenia.Context(0)
    yield c
    c.close()


def run_alone(ctx, desc, seed, steps, init=None):
    import lenia

    with lenia.World(ctx, desc) as w:
        w.reset(seed)
        if init is not None:
            w.write(init)
        w.frame(0, 1)
        w.step(steps)
        return w.read()


def run_batched(ctx, descs, seeds, steps, graph_steps, inits=None, grids=None):
    import lenia

    worlds = [lenia.World(ctx, d) for d in descs]
    try:
        for i, (w, s) in enumerate(zip(worlds, seeds)):
            w.reset(s)
            if inits is not None:
                w.write(inits[i])
            w.frame(0, 1)
        with lenia.Batch(ctx, worlds) as b:
            if grids is None:
                b.step(steps, graph_steps)
            else:
                # several launch sequences, each with its own grids: the tile
                # queues must start every launch from zero
                done = 0
                for g in grids:
                    b.set_grids(*g)
                    assert b.grids() == tuple(g)
                    n = steps // len(grids)
                    b.step(n, graph_steps if graph_steps and n % graph_steps == 0 else 0)
                    done += n
                b.step(steps - done)
        return [w.read() for w in worlds]
    finally:
        for w in worlds:
            w.close()


@pytest.mark.parametrize("graph_steps", [0, 8])
def test_fft_batch_bit_exact_vs_alone(ctx, graph_steps):
    descs = sweep_worlds(24, seed=7)
    seeds = [1000 + i for i in range(len(descs))]
    got = run_batched(ctx, descs, seeds, 24, graph_steps)
    for i, d in enumerate(descs):
        want = run_alone(ctx, d, seeds[i], 24)
        assert got[i].tobytes() == want.tobytes(), (i, d["width"], d["height"])
        print(f"world {i} {d['width']}x{d['height']} K={len(d['kernels'])}: bit-exact, "
              f"mass {float(want.sum()):.3f}")


QUEUES_WITHDRAWN = ("tile-queue grids withdrawn: this test aborted on the GPU (2026-12-04) once the "
                    "column pass ran several launches side by side; the cause was not found, and the "
                    "queues, 3x slower than one workgroup per tile, were removed (lenia_batch_set_grids "
                    "now refuses non-zero grids)")


@pytest.mark.skip(reason=QUEUES_WITHDRAWN)
@pytest.mark.parametrize("grids", [[(1, 1)], [(2, 4), (1, 2)], [(4, 1), (0, 0), (1, 0)]])
def test_fft_batch_tile_queues_bit_exact(ctx, grids):
    """resident grids taking tiles from the queues (a few workgroups per CU,
    so each takes many tiles of several worlds), switched between launch
    sequences and replayed from graphs: the same bytes as each world alone"""
    descs = sweep_worlds(24, seed=7)
    seeds = [1000 + i for i in range(len(descs))]
    got = run_batched(ctx, descs, seeds, 24, 4, grids=grids)
    for i, d in enumerate(descs):
        want = run_alone(ctx, d, seeds[i], 24)
        assert got[i].tobytes() == want.tobytes(), (grids, i, d["width"], d["height"])


@pytest.mark.skip(reason=QUEUES_WITHDRAWN)
def test_fft_batch_tile_queues_many_worlds(ctx):
    """the 256-world sweep with one resident workgroup per CU per pass (more
    than ten tiles per workgroup), two graph replays"""
    descs = sweep_worlds(256)
    seeds = [1000 + i for i in range(256)]
    got = run_batched(ctx, descs, seeds, 16, 8, grids=[(1, 1)])
    for i in range(0, 256, 5):
        want = run_alone(ctx, descs[i], seeds[i], 16)
        assert got[i].tobytes() == want.tobytes(), (i, descs[i]["width"], descs[i]["height"])


def test_direct_batch_bit_exact_vs_alone(ctx):
    descs = []
    for i, (W, H, R) in enumerate([(64, 64, 8), (96, 80, 10), (130, 70, 13), (200, 128, 20)]):
        descs.append(dict(width=W, height=H, convMode=DIRECT, growthMode=0, dt=0.1, radius=R,
                          kernels=[ring(R, [1, 0.5][: 1 + i % 2], 0.15 + 0.02 * i, 0.02, 1.0)],
                          initMode=0, initRect=(W // 4, H // 4, W // 2, H // 2)))
    seeds = [5, 6, 7, 8]
    got = run_batched(ctx, descs, seeds, 17, 0)
    for i, d in enumerate(descs):
        assert got[i].tobytes() == run_alone(ctx, d, seeds[i], 17).tobytes(), i


def test_direct_batch_of_many_tiles_bit_exact_vs_alone(ctx):
    """40 direct worlds of 128x128 (8 output tiles each): the batch's 320
    tiles run the 256-thread form of the FP32 direct kernel, each world alone
    (8 tiles) the 512-thread form that splits each output's window rows over
    two threads; both must give the same bytes."""
    descs = [dict(width=128, height=128, convMode=DIRECT, growthMode=i % 2, dt=0.1, radius=9 + i % 5,
                  kernels=[ring(9 + i % 5, [1, 0.5][: 1 + i % 2], 0.15 + 0.002 * i, 0.02, 1.0)],
                  initMode=0, initRect=(32, 32, 64, 64)) for i in range(40)]
    seeds = [300 + i for i in range(40)]
    got = run_batched(ctx, descs, seeds, 9, 0)
    for i in range(0, 40, 3):
        assert got[i].tobytes() == run_alone(ctx, descs[i], seeds[i], 9).tobytes(), i


def test_batch_with_extras_bit_exact(ctx):
    base = dict(width=96, height=64, convMode=FFT, growthMode=1, dt=0.35, radius=8,
                noiseAmount=0.4, muJitter=0.5, injectAmount=0.3, injectCount=3,
                kernels=[ring(8, [1], 0.15, 0.017, 1.0), ring(6, [1, 0.6], 0.3, 0.035, 0.6)],
                initMode=0, initRect=(0, 0, 0, 0))
    plain = dict(base, noiseAmount=0.0, muJitter=0.0, injectAmount=0.0)
    descs = [base, plain, dict(base, width=128)]
    seeds = [11, 12, 13]
    got = run_batched(ctx, descs, seeds, 16, 8)
    for i, d in enumerate(descs):
        assert got[i].tobytes() == run_alone(ctx, d, seeds[i], 16).tobytes(), i


def test_batch_sees_parameter_changes(ctx):
    import lenia

    d = sweep_worlds(4, seed=3)
    worlds = [lenia.World(ctx, x) for x in d]
    try:
        for i, w in enumerate(worlds):
            w.reset(i)
            w.frame(0, 1)
        with lenia.Batch(ctx, worlds) as b:
            b.step(4)
            d2 = dict(d[0], dt=d[0]["dt"] / 2)
            worlds[0].set_params(d2)
            b.step(4)
            got = worlds[0].read()
        with lenia.World(ctx, d[0]) as w:
            w.reset(0)
            w.frame(0, 1)
            w.step(4)
            w.set_params(d2)
            w.step(4)
            assert got.tobytes() == w.read().tobytes()
    finally:
        for w in worlds:
            w.close()


def test_batch_errors(ctx):
    import lenia

    a = lenia.World(ctx, sweep_worlds(1, seed=1)[0])
    b = lenia.World(ctx, dict(width=64, height=64, convMode=DIRECT, radius=5,
                              kernels=[exp_shell(5)], initMode=INIT_NONE))
    try:
        with pytest.raises(lenia.LeniaError):
            lenia.Batch(ctx, [a, b])  # mixed forms
        with pytest.raises(lenia.LeniaError):
            lenia.Batch(ctx, [a, a])  # duplicate
        with lenia.Batch(ctx, [a]) as bt:
            with pytest.raises(lenia.LeniaError):
                bt.step(10, 4)  # not a multiple of the graph length
            with pytest.raises(lenia.LeniaError):
                bt.set_grids(65, 0)  # tile-queue grids are withdrawn: non-zero grids refused
        with lenia.Batch(ctx, [b]) as bt:
            with pytest.raises(lenia.LeniaError):
                bt.set_grids(1, 1)  # refused for every form
            bt.set_grids(0, 0)  # one workgroup per tile: accepted
            assert bt.grids() == (0, 0)
    finally:
        a.close()
        b.close()


def test_sweep_256_one_step_matches_oracle_sample(ctx):
    """the full 256-world sweep in one batch; a sample of worlds checked against
    the FP64 oracle after one step from a random state"""
    import lenia
    import oracle as O

    descs = [dict(d, initMode=INIT_NONE) for d in sweep_worlds(256)]
    worlds = [lenia.World(ctx, d) for d in descs]
    try:
        states = [rand_state(d["width"], d["height"], seed=i) for i, d in enumerate(descs)]
        for w, A in zip(worlds, states):
            w.write(A)
        with lenia.Batch(ctx, worlds) as bt:
            bt.step(1)
        for i in range(0, 256, 37):
            want = O.step(descs[i], states[i].astype(np.float64), use_fft=True)
            mx = float(np.abs(worlds[i].read().astype(np.float64) - want).max())
            print(f"sweep world {i} {descs[i]['width']}x{descs[i]['height']}: max-abs {mx:.2e}")
            assert mx <= 1e-5
    finally:
        for w in worlds:
            w.close()


def test_sweep_256_as_benchmarked(ctx):
    """The batched sweep exactly as it is benchmarked: all 256 worlds, launch
    shapes from lenia_batch_autotune, 50-step graph replay, 250 steps (one
    50-step warm-up sequence and 200 more).  Every world is byte-compared with
    the same world stepped alone (its own default launch shapes and 32-step
    graphs); then one more batched step from the final states is checked
    against the FP64 oracle on a sample that favours the worlds still alive."""
    import lenia
    import oracle as O
    from lenia_cases import live_cells

    descs = sweep_worlds(256)
    seeds = [1000 + i for i in range(256)]
    worlds = [lenia.World(ctx, d) for d in descs]
    try:
        for w, s in zip(worlds, seeds):
            w.reset(s)
            w.frame(0, 1)
        with lenia.Batch(ctx, worlds) as bt:
            ms = bt.autotune(10)
            print(f"sweep 256: tuned executors (row, column waves per workgroup) {bt.waves()}")
            bt.step(50, 50)
            bt.step(200, 50)
            got = [w.read() for w in worlds]
            alive = [i for i, A in enumerate(got) if live_cells(A) > 0]
            print(f"sweep 256: tuned {ms * 1e3:.1f} us per batched step; {len(alive)} of 256 "
                  f"worlds alive after 250 steps")
            for i, d in enumerate(descs):
                with lenia.World(ctx, d) as w:
                    w.reset(seeds[i])
                    w.frame(0, 1)
                    w.step(250)
                    assert got[i].tobytes() == w.read().tobytes(), (i, d["width"], d["height"])
            assert len(alive) >= 8
            sample = alive[:: max(1, len(alive) // 8)][:8] + list(range(3, 256, 64))
            bt.step(1)
            for i in sample:
                want = O.step(descs[i], got[i].astype(np.float64), use_fft=True)
                mx = float(np.abs(worlds[i].read().astype(np.float64) - want).max())
                print(f"sweep world {i} {descs[i]['width']}x{descs[i]['height']} "
                      f"K={len(descs[i]['kernels'])} ({live_cells(got[i])} live): step 251 "
                      f"max-abs {mx:.2e}")
                assert mx <= 1e-5
    finally:
        for w in worlds:
            w.close()
