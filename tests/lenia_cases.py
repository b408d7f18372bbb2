"""World descriptions shared by the GPU tests, the smoke entry point and the
benchmark.  A world is a plain dict understood by both the library binding
(lenia._abi.world_desc) and the oracle (oracle.world_struct)."""
from __future__ import annotations

import numpy as np

SHELL, RING = 0, 1
GAUSS, EXP = 0, 1
DIRECT, FFT, BF16, BF16X3 = 0, 1, 2, 3
INIT_PATCH, INIT_NOISE, INIT_NONE = 0, 1, 2


def exp_shell(R, mu=0.15, sigma=0.015, weight=1.0):
    """exponential-bump one-shell kernel exp(4 - 1/(r(1-r)))"""
    return dict(shape=SHELL, core=EXP, radius=float(R), shellMu=0.5, shellSigma=0.15,
                betas=[1.0], mu=mu, sigma=sigma, weight=weight)


def gauss_shell(R, mu=0.15, sigma=0.015, weight=1.0, shellMu=0.5, shellSigma=0.15):
    """the reference's buildKernelTexture kernel"""
    return dict(shape=SHELL, core=GAUSS, radius=float(R), shellMu=shellMu,
                shellSigma=shellSigma, betas=[1.0], mu=mu, sigma=sigma, weight=weight)


def ring(R, betas, mu, sigma, weight, core=GAUSS, shellSigma=0.15):
    """the reference's buildRingKernelImage kernel"""
    return dict(shape=RING, core=core, radius=float(R), shellMu=0.5, shellSigma=shellSigma,
                betas=list(betas), mu=mu, sigma=sigma, weight=weight)


def centred(W, H, pw, ph):
    return ((W - pw) // 2, (H - ph) // 2, pw, ph)


def config0(conv=DIRECT, core=EXP):
    """512x512, R=13 one-shell kernel, mu 0.15 sigma 0.015, dt 0.1, 128^2 patch"""
    k = exp_shell(13) if core == EXP else gauss_shell(13)
    return dict(width=512, height=512, convMode=conv, growthMode=0, dt=0.1, radius=13,
                kernels=[k], initMode=INIT_PATCH, initRect=centred(512, 512, 128, 128))


C1_KERNELS = [ring(24, [1.0], 0.15, 0.015, 0.5),
              ring(24, [1.0, 0.5], 0.26, 0.036, 0.3),
              ring(24, [0.5, 1.0, 0.25], 0.12, 0.012, 0.2)]


def config1(W=1024, H=512, conv=FFT):
    """three ring kernels R=24, betas [1], [1,.5], [.5,1,.25], w .5/.3/.2"""
    return dict(width=W, height=H, convMode=conv, growthMode=0, dt=0.1, radius=24,
                kernels=[dict(k) for k in C1_KERNELS], initMode=INIT_PATCH,
                initRect=centred(W, H, 256, 256))


def config2(conv=DIRECT):
    """1280x720, one exp-bump kernel R=32, full-frame noise"""
    return dict(width=1280, height=720, convMode=conv, growthMode=0, dt=0.1, radius=32,
                kernels=[exp_shell(32)], initMode=INIT_PATCH, initRect=(0, 0, 0, 0))


def config4(conv=FFT):
    """the differentiable-rollout world: 256x256, two ring kernels R=16 with
    growth (mu, sigma) = (0.30, 0.05) and (0.40, 0.07), weights .6/.4, full-frame
    seeded noise; from noise the field keeps 60-70% of its cells strictly
    inside (0, 1) for the first 20 steps, so gradients flow through most cells
    (the sharper (0.15, 0.017) growth of the one-kernel worlds empties a noise
    field within 20 steps)"""
    return dict(width=256, height=256, convMode=conv, growthMode=0, dt=0.1, radius=16,
                kernels=[ring(16, [1.0], 0.30, 0.05, 0.6),
                         ring(16, [1.0, 0.5], 0.40, 0.07, 0.4)],
                initMode=INIT_PATCH, initRect=(0, 0, 256, 256))


SWEEP_SIDES = [64, 96, 128, 192, 256, 384, 512]


def sweep_worlds(n=256, seed=20260709, conv=FFT):
    """the batched sweep: per-world sides, R, mu, sigma, dt, 1-3 kernels with
    random shell peaks, drawn from a SplitMix64 stream (LifeCore/Math/Random.h)"""
    state = [(seed + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF]

    def nxt():
        state[0] = (state[0] + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def f01():
        return (nxt() >> 40) / 16777216.0

    worlds = []
    for _ in range(n):
        W = SWEEP_SIDES[nxt() % len(SWEEP_SIDES)]
        H = SWEEP_SIDES[nxt() % len(SWEEP_SIDES)]
        R = 8 + nxt() % 17
        R = min(R, (min(W, H) - 1) // 2)
        mu = 0.10 + 0.20 * f01()
        sigma = 0.01 + 0.04 * f01()
        dt = [0.05, 0.1, 0.2][nxt() % 3]
        K = 1 + nxt() % 3
        ks = []
        for k in range(K):
            nb = 1 + nxt() % 3
            betas = [max(f01(), 0.05) for _ in range(nb)]
            ks.append(ring(R, betas, mu, sigma, 1.0 / K))
        pw, ph = W // 2, H // 2
        worlds.append(dict(width=W, height=H, convMode=conv, growthMode=0, dt=dt, radius=R,
                           kernels=ks, initMode=INIT_PATCH, initRect=centred(W, H, pw, ph)))
    return worlds


def live_cells(A) -> int:
    """cells strictly inside (0, 1): the ones whose growth is not clipped away"""
    A = np.asarray(A)
    return int(np.count_nonzero((A > 0) & (A < 1)))


def assert_alive(A, prev=None, min_live=256, min_std=0.0, min_change=0.0, label=""):
    """a multi-step comparison means something only on a field that is still
    evolving: at least `min_live` cells strictly inside (0, 1), a spatial
    standard deviation of at least `min_std` (not a uniform fixed point) and,
    given the previous state, a largest one-step change of at least
    `min_change`"""
    A = np.asarray(A, np.float64)
    n = live_cells(A)
    assert n >= min_live, f"{label}: field is dead ({n} live cells < {min_live})"
    sd = float(A.std())
    assert sd >= min_std, f"{label}: field is uniform (std {sd:.3e} < {min_std})"
    if prev is not None:
        ch = float(np.abs(A - np.asarray(prev, np.float64)).max())
        assert ch >= min_change, f"{label}: field is static (max change {ch:.3e} < {min_change})"
    return n


def config1_live(W=1024, H=512, conv=FFT):
    """the headline world's three R=24 ring kernels (same shells, weights and
    growth) in growth mode 1, for trajectories that stay alive: started from a
    value-noise field (coverage 0.42, scale 26) it keeps every cell inside
    (0, 1) with a growing spatial spread (std 0.09 -> 0.15 over 100 steps) and
    per-step changes of ~0.05, where the mode-0 patch world is empty by step
    18"""
    return dict(config1(W, H, conv), growthMode=1, initMode=INIT_NONE)


def rand_state(W, H, seed=0, zero_frac=0.3):
    rng = np.random.default_rng(seed)
    A = rng.random((H, W)).astype(np.float32)
    A[rng.random((H, W)) < zero_frac] = 0.0
    return A
