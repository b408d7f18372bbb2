"""Writes tests/golden/lenia_golden.npz: FP64 golden vectors for the Lenia path.

Independent of the oracle: every quantity here is computed with vectorised
numpy (np.roll gathers, numpy.fft) and, for the rollout gradient, torch
float64 autograd -- not with the oracle's C code.  The formulas follow the
reference's own definitions:

  hashes       Shaders/Common/Random.metal:5-19 (uint32 arithmetic in numpy)
  kernels      Modules/FieldModules/Lenia.cpp:95-125 (ring image) and
               :410-436 (shell window)
  step         Shaders/Field/Lenia.metal:63-128 (direct) and :161-232 (growth)
  FFT          numpy.fft (forward e^{-i...}, unscaled)

Run:  python tests/golden/make_golden.py   (writes next to this file)
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
U32 = np.uint32


# ------------------------------------------------------------------ hashes
def pcg(v):
    v = np.asarray(v, dtype=np.uint64) & 0xFFFFFFFF
    state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return ((word >> 22) ^ word) & 0xFFFFFFFF


def rand01(x, y, tag, seed):
    t = pcg(np.uint64(tag) ^ np.uint64(seed))
    h = pcg(np.asarray(x, dtype=np.uint64) ^ pcg(np.asarray(y, dtype=np.uint64) ^ t))
    return (h >> 8).astype(np.float64) / 16777216.0


# ------------------------------------------------------------------ kernels
def core(kind, x, shell_mu, sigma):
    if kind == "exp":
        out = np.zeros_like(x)
        m = (x > 0) & (x < 1)
        out[m] = np.exp(4.0 - 1.0 / (x[m] * (1.0 - x[m])))
        return out
    return np.exp(-0.5 * ((x - shell_mu) / sigma) ** 2)


def raw_kernel(k, dx, dy):
    rr = np.hypot(dx, dy)
    if k["shape"] == "shell":
        r = rr / k["radius"]
        val = core(k["core"], r, k["shellMu"], k["shellSigma"])
        return np.where((r > 0) & (r <= 1), val, 0.0)
    betas = np.asarray(k["betas"], dtype=np.float64)
    B = len(betas)
    u = rr / k["radius"] * B
    i = np.minimum(np.floor(u).astype(int), B - 1)
    frac = u - i
    if k["core"] == "exp":
        c = core("exp", frac, 0.5, 1.0)
    else:
        c = np.exp(-0.5 * ((frac - 0.5) / (k["shellSigma"] * B)) ** 2)
    val = betas[np.clip(i, 0, B - 1)] * c
    return np.where((rr > 0) & (rr <= k["radius"]), val, 0.0)


def window(k):
    h = int(np.ceil(k["radius"] - 1e-9))
    d = np.arange(-h, h + 1, dtype=np.float64)
    dy, dx = np.meshgrid(d, d, indexing="ij")
    w = raw_kernel(k, dx, dy)
    return w / w.sum()


def image(k, W, H):
    x = np.arange(W)
    y = np.arange(H)
    dx = np.where(x > W // 2, x - W, x).astype(np.float64)
    dy = np.where(y > H // 2, y - H, y).astype(np.float64)
    DY, DX = np.meshgrid(dy, dx, indexing="ij")
    w = raw_kernel(k, DX, DY)
    return w / w.sum()


# ------------------------------------------------------------------ step
def potential_roll(A, win):
    """u(y,x) = sum K[dy][dx] A[y+dy][x+dx] by whole-array rolls."""
    h = (win.shape[0] - 1) // 2
    u = np.zeros_like(A)
    for iy in range(win.shape[0]):
        for ix in range(win.shape[1]):
            w = win[iy, ix]
            if w != 0.0:
                u += w * np.roll(A, shift=(-(iy - h), -(ix - h)), axis=(0, 1))
    return u


def potential_npfft(A, img):
    return np.real(np.fft.ifft2(np.fft.fft2(A) * np.fft.fft2(img)))


def growth(world, A, us, seed, frame):
    H, W = A.shape
    y, x = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    muJ = 0.0
    if world.get("muJitter", 0.0) > 0:
        muJ = (rand01(x, y, 401 + frame % 613, seed) - 0.5) * world["muJitter"] * 0.05
    sw = swb = sstd = 0.0
    for k, u in zip(world["kernels"], us):
        bell = np.exp(-0.5 * ((u - (k["mu"] + muJ)) / k["sigma"]) ** 2)
        swb = swb + k["weight"] * bell
        sw = sw + k["weight"]
        sstd = sstd + k["weight"] * (2 * bell - 1)
    if world.get("growthMode", 0) == 1:
        target = swb / sw if sw > 1e-6 else A
        A2 = A + world["dt"] * (target - A)
    else:
        A2 = A + world["dt"] * sstd
    if world.get("noiseAmount", 0.0) > 0:
        A2 = A2 + (rand01(x, y, 503, seed) - 0.5) * world["noiseAmount"] * 0.03
    if world.get("injectAmount", 0.0) > 0:
        uvx, uvy = (x + 0.5) / W, (y + 0.5) / H
        aspect = W / H
        cycle = frame // 12
        blobR = world["radius"] * 0.9 / H
        for i in range(world.get("injectCount", 3)):
            cx = rand01(i, cycle, 811, seed)
            cy = rand01(cycle, i, 823, seed)
            dx = uvx - cx
            dy = uvy - cy
            dx = dx - np.sign(dx) * np.floor(np.abs(dx) + 0.5)
            dy = dy - np.sign(dy) * np.floor(np.abs(dy) + 0.5)
            dx = dx * aspect
            r2 = (dx * dx + dy * dy) / (blobR * blobR)
            t = np.clip(np.exp(-0.5 * r2) * world["injectAmount"], 0, 1)
            A2 = A2 + (0.6 - A2) * t
    return np.clip(A2, 0.0, 1.0)


def step(world, A, use_fft, seed=0, frame=0):
    H, W = A.shape
    us = []
    for k in world["kernels"]:
        if use_fft:
            us.append(potential_npfft(A, image(k, W, H)))
        else:
            us.append(potential_roll(A, window(k)))
    return growth(world, A, us, seed, frame)


# ------------------------------------------------------------------ gradient (torch f64)
def torch_rollout_grad(world, A0, T, gT):
    import torch

    H, W = A0.shape
    dt = torch.tensor(world["dt"], dtype=torch.float64, requires_grad=True)
    A = torch.tensor(A0, dtype=torch.float64, requires_grad=True)
    params = []
    specs = []
    x = np.arange(W)
    y = np.arange(H)
    dxs = np.where(x > W // 2, x - W, x).astype(np.float64)
    dys = np.where(y > H // 2, y - H, y).astype(np.float64)
    DY, DX = np.meshgrid(dys, dxs, indexing="ij")
    for k in world["kernels"]:
        mu = torch.tensor(k["mu"], dtype=torch.float64, requires_grad=True)
        sg = torch.tensor(k["sigma"], dtype=torch.float64, requires_grad=True)
        wk = torch.tensor(k["weight"], dtype=torch.float64, requires_grad=True)
        betas = torch.tensor(k["betas"], dtype=torch.float64, requires_grad=True)
        B = len(k["betas"])
        shells = []
        for j in range(B):
            kk = dict(k)
            kk["betas"] = [1.0 if q == j else 0.0 for q in range(B)]
            shells.append(torch.tensor(raw_kernel(kk, DX, DY)))
        raw = sum(betas[j] * shells[j] for j in range(B))
        img = raw / raw.sum()
        specs.append(torch.fft.fft2(img))
        params.append((mu, sg, wk, betas))
    cur = A
    for _ in range(T):
        FA = torch.fft.fft2(cur)
        s = 0.0
        for (mu, sg, wk, _b), spec in zip(params, specs):
            u = torch.real(torch.fft.ifft2(FA * spec))
            bell = torch.exp(-0.5 * ((u - mu) / sg) ** 2)
            s = s + wk * (2 * bell - 1)
        cur = torch.clamp(cur + dt * s, 0.0, 1.0)
    loss = (cur * torch.tensor(gT)).sum()
    leaves = [A, dt] + [p for ps in params for p in ps]
    grads = torch.autograd.grad(loss, leaves)
    out = {"A0": grads[0].numpy(), "dt": float(grads[1])}
    K = len(params)
    out["mu"] = np.array([float(grads[2 + 4 * i]) for i in range(K)])
    out["sigma"] = np.array([float(grads[3 + 4 * i]) for i in range(K)])
    out["weight"] = np.array([float(grads[4 + 4 * i]) for i in range(K)])
    out["betas"] = [grads[5 + 4 * i].numpy() for i in range(K)]
    out["AT"] = cur.detach().numpy()
    return out


# ------------------------------------------------------------------ fixtures
KERNELS = {
    "shell_gauss_R5": dict(shape="shell", core="gaussian", radius=5.0, shellMu=0.5,
                           shellSigma=0.15, betas=[1.0]),
    "shell_gauss_R7_mu04": dict(shape="shell", core="gaussian", radius=7.0, shellMu=0.4,
                                shellSigma=0.2, betas=[1.0]),
    "shell_exp_R6": dict(shape="shell", core="exp", radius=6.0, shellMu=0.5, shellSigma=0.15,
                         betas=[1.0]),
    "ring_b1_R5": dict(shape="ring", core="gaussian", radius=5.0, shellSigma=0.15, betas=[1.0]),
    "ring_b2_R6p5": dict(shape="ring", core="gaussian", radius=6.5, shellSigma=0.15,
                         betas=[1.0, 0.5]),
    "ring_b3_R7_exp": dict(shape="ring", core="exp", radius=7.0, shellSigma=0.15,
                           betas=[0.5, 1.0, 0.25]),
}


def kdict(name, **growth_params):
    k = dict(KERNELS[name])
    k.update(growth_params)
    return k


WORLDS = {
    # single shell kernel, direct, growth mode 0 (the leniaStep formula)
    "w1_direct": dict(width=32, height=24, growthMode=0, dt=0.1, radius=5, kernels=[
        kdict("shell_gauss_R5", mu=0.15, sigma=0.015, weight=1.0)]),
    # exp core, mode 0
    "w1_exp": dict(width=30, height=26, growthMode=0, dt=0.1, radius=6, kernels=[
        kdict("shell_exp_R6", mu=0.15, sigma=0.015, weight=1.0)]),
    # three ring kernels, mode 0, FFT and direct
    "w3_mode0": dict(width=36, height=28, growthMode=0, dt=0.1, radius=7, kernels=[
        kdict("ring_b1_R5", mu=0.15, sigma=0.015, weight=0.5),
        kdict("ring_b2_R6p5", mu=0.26, sigma=0.036, weight=0.3),
        kdict("ring_b3_R7_exp", mu=0.12, sigma=0.012, weight=0.2)]),
    # mode 1 with noise, jitter and inject (the preset terms)
    "w2_mode1_extras": dict(width=40, height=24, growthMode=1, dt=0.35, radius=6,
                            noiseAmount=0.5, muJitter=0.8, injectAmount=0.4, injectCount=3,
                            kernels=[kdict("ring_b1_R5", mu=0.15, sigma=0.017, weight=1.0),
                                     kdict("ring_b2_R6p5", mu=0.3, sigma=0.035, weight=0.6)]),
}


def main():
    rng = np.random.default_rng(20260709)
    out = {}
    # hashes
    v = np.concatenate([np.arange(16, dtype=np.uint64),
                        rng.integers(0, 2**32, size=48, dtype=np.uint64)])
    out["hash_in"] = v.astype(np.uint32)
    out["hash_out"] = pcg(v).astype(np.uint32)
    xs = rng.integers(0, 4096, size=32)
    ys = rng.integers(0, 4096, size=32)
    out["rand_xy"] = np.stack([xs, ys]).astype(np.uint32)
    out["rand_out"] = rand01(xs, ys, 503, 3628426083)
    # kernels
    for name, k in KERNELS.items():
        out["win_" + name] = window(k)
        if k["shape"] == "ring":
            out["img_" + name] = image(k, 24, 20)
    # FFT spectra
    for n in (15, 60, 64, 96, 120):
        x = rng.random(n) + 1j * rng.random(n)
        out[f"fft_in_{n}"] = x
        out[f"fft_out_{n}"] = np.fft.fft(x)
    X = rng.random((12, 30))
    out["fft2_in"] = X
    out["fft2_out"] = np.fft.fft2(X)
    # steps
    for name, w in WORLDS.items():
        H, W = w["height"], w["width"]
        A = rng.random((H, W))
        A[rng.random((H, W)) < 0.3] = 0.0
        out[f"step_{name}_A"] = A
        out[f"step_{name}_direct"] = step(w, A, use_fft=False, seed=2831084092, frame=29)
        out[f"step_{name}_fft"] = step(w, A, use_fft=True, seed=2831084092, frame=29)
    # rollout gradient (torch float64 autograd)
    w = dict(width=24, height=20, growthMode=0, dt=0.1, radius=4, kernels=[
        dict(shape="ring", core="gaussian", radius=4.0, shellSigma=0.15, betas=[1.0],
             mu=0.2, sigma=0.05, weight=0.6),
        dict(shape="ring", core="gaussian", radius=5.0, shellSigma=0.15, betas=[1.0, 0.6],
             mu=0.25, sigma=0.06, weight=0.4)])
    A0 = 0.2 + 0.6 * rng.random((20, 24))
    gT = rng.standard_normal((20, 24))
    g = torch_rollout_grad(w, A0, 6, gT)
    out["grad_A0"] = A0
    out["grad_gT"] = gT
    out["grad_AT"] = g["AT"]
    out["grad_dA0"] = g["A0"]
    out["grad_dt"] = np.array([g["dt"]])
    out["grad_mu"] = g["mu"]
    out["grad_sigma"] = g["sigma"]
    out["grad_weight"] = g["weight"]
    out["grad_betas0"] = g["betas"][0]
    out["grad_betas1"] = g["betas"][1]
    np.savez_compressed(os.path.join(HERE, "lenia_golden.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
