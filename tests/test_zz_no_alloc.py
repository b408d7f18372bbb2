"""Allocation-free stepping (lenia_fft_step_scratch) through the C ABI.

The scratch-backed step must give the same bytes as lenia_world_step on a
world of the same descriptor and seed holding the same state, on two calls in
a row with a smaller field the second time (and back again), with and without
the frame-seeded noise / jitter / inject terms; it must not allocate (the
context's device-memory accounting is unchanged) and must reject fields whose
buffers exceed its 64 MiB array, non-FFT forms and NULL arguments."""
import ctypes as C

import numpy as np
import pytest

from lenia_cases import FFT, INIT_NONE, config1, config4, rand_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import lenia

    c = lenia.Context(0)
    yield c
    c.close()


def scratch_step(ctx, desc, seed, frame, ptr, n):
    from lenia import _abi

    d = _abi.world_desc(desc)
    return _abi.lib().lenia_fft_step_scratch(ctx.handle, C.byref(d), seed, frame,
                                             C.c_void_p(ptr), n)


def world_reference(ctx, desc, seed, frame, A, n):
    """lenia_world_step's result: a world reset with `seed` holding A, one
    frame at `frame` then n - 1 more steps (frames frame+1 ...)"""
    import lenia

    with lenia.World(ctx, desc) as w:
        w.reset(seed)
        w.write(A)
        w.frame(frame, 1)
        if n > 1:
            w.step(n - 1)
        return w.read()


CASES = [
    ("1024x512_K3", dict(config1(), initMode=INIT_NONE), 7777, 0, 40),
    ("256x256_K2_extras_smaller", dict(config4(), initMode=INIT_NONE, noiseAmount=0.4,
                                       muJitter=0.5, injectAmount=0.3, injectCount=3),
     4242, 29, 37),
    ("1024x512_K3_again", dict(config1(), initMode=INIT_NONE, growthMode=1), 99, 5, 33),
]


def test_scratch_step_equals_world_step(ctx):
    import lenia

    for name, desc, seed, frame, n in CASES:
        A = rand_state(desc["width"], desc["height"], seed=len(name))
        want = world_reference(ctx, desc, seed, frame, A, n)
        with lenia.World(ctx, desc) as holder:  # owns the device field only
            holder.write(A)
            ptr = holder.state_device_ptr()
            ctx.sync()
            before = ctx.memory()
            rc = scratch_step(ctx, desc, seed, frame, ptr, n)
            assert rc == 0, (name, rc, lenia._abi.lib().lenia_last_error())
            ctx.sync()
            assert ctx.memory() == before, name
            got = holder.read()
        same = got.tobytes() == want.tobytes()
        print(f"{name}: {n} steps from frame {frame}: scratch == world step: {same} "
              f"(max-abs {float(np.abs(got - want).max()):.2e})")
        assert same, name


def test_scratch_step_rejects_what_it_cannot_serve(ctx):
    from lenia import _abi

    L = _abi.lib()
    with __import__("lenia").World(ctx, dict(config4(), initMode=INIT_NONE)) as holder:
        ptr = holder.state_device_ptr()
        big = dict(config1(4096, 4096), initMode=INIT_NONE)
        assert scratch_step(ctx, big, 1, 0, ptr, 1) == _abi.ERR_INVALID_ARG
        msg = L.lenia_last_error().decode()
        print(msg)
        assert "maximum" in msg and "4096x4096" in msg
        direct = dict(config4(), convMode=0, initMode=INIT_NONE)
        assert scratch_step(ctx, direct, 1, 0, ptr, 1) == -2
        assert scratch_step(ctx, config4(), 1, 0, 0, 1) == _abi.ERR_INVALID_ARG
        assert L.lenia_fft_step_scratch(None, None, 0, 0, None, 1) == _abi.ERR_INVALID_ARG
