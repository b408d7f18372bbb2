"""torch.autograd binding of the differentiable Lenia rollout.

The forward and reverse passes both run in liblenia_mi355x.so (C ABI
lenia_rollout_forward / lenia_rollout_backward); torch only carries the
tensors and the optimiser.  Growth mode 0, FFT form, no noise / jitter /
inject (the terms the reverse pass differentiates).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from ._abi import check

if _abi._lib is not None and not _abi.loaded_after_torch:
    raise ImportError("liblenia_mi355x.so was loaded before torch: import torch (or "
                      "lenia.autograd) before the first library call so that both use "
                      "one HIP runtime")


def _world_desc_with(world, mu, sigma, weight, betas, dt) -> dict:
    d = dict(world.desc)
    d["kernels"] = [dict(k) for k in world.desc["kernels"]]
    d["dt"] = float(dt)
    for k, kd in enumerate(d["kernels"]):
        kd["mu"] = float(mu[k])
        kd["sigma"] = float(sigma[k])
        kd["weight"] = float(weight[k])
        nb = len(kd.get("betas", [1.0]))
        kd["betas"] = [float(betas[k][j]) for j in range(nb)]
    return d


def _apply(world, desc: dict) -> None:
    if desc != world.desc:
        world.set_params(desc)


def _sync_torch():
    if torch.cuda.is_available():
        torch.cuda.current_stream().synchronize()


class LeniaRollout(torch.autograd.Function):
    """A_T = rollout(A0; mu, sigma, weight, betas, dt) over T steps of `world`.

    mu, sigma, weight: (K,) float64 tensors; betas: (K, 4) float64 (unused
    shells ignored); dt: 0-d float64 tensor; A0: (H, W) float32 CUDA tensor.
    """

    @staticmethod
    def forward(ctx, A0, mu, sigma, weight, betas, dt, world, T: int, ckpt_every: int):
        _apply(world, _world_desc_with(world, mu, sigma, weight, betas, dt))
        A0c = A0.detach().contiguous()
        AT = torch.empty_like(A0c)
        _sync_torch()
        world._rollouts += 1
        check(_abi.lib().lenia_rollout_forward(world.handle, T, ckpt_every, A0c.data_ptr(),
                                               AT.data_ptr(), 1))
        ctx.save_for_backward(A0c, mu, sigma, weight, betas, dt)
        ctx.world, ctx.T, ctx.ckpt = world, T, ckpt_every
        ctx.world_version = world.version_token()
        return AT

    @staticmethod
    def backward(ctx, gAT):
        A0c, mu, sigma, weight, betas, dt = ctx.saved_tensors
        world = ctx.world
        _apply(world, _world_desc_with(world, mu, sigma, weight, betas, dt))
        g = gAT.detach().to(torch.float32).contiguous()
        gA0 = torch.empty_like(A0c)
        pg = _abi.ParamGrads()
        _sync_torch()
        # the checkpoints of this forward are reused when nothing changed since
        # (same world, T, spacing and parameters); else the sweep reruns from A0
        ws_ok = ctx.world_version == world.version_token()
        if not ws_ok:
            world._rollouts += 1
        check(_abi.lib().lenia_rollout_backward(world.handle, ctx.T, ctx.ckpt,
                                                None if ws_ok else A0c.data_ptr(),
                                                g.data_ptr(), gA0.data_ptr(), C.byref(pg), 1))
        K = mu.shape[0]
        dev = mu.device
        f64 = torch.float64
        dmu = torch.tensor(pg.dMu[:K], dtype=f64, device=dev)
        dsg = torch.tensor(pg.dSigma[:K], dtype=f64, device=dev)
        dw = torch.tensor(pg.dWeight[:K], dtype=f64, device=dev)
        db = torch.tensor([[pg.dBetas[k][j] for j in range(betas.shape[1])] for k in range(K)],
                          dtype=f64, device=dev)
        ddt = torch.tensor(pg.dDt, dtype=f64, device=dev)
        return gA0, dmu, dsg, dw, db, ddt, None, None, None


def rollout(world, A0, mu, sigma, weight, betas, dt, T: int, ckpt_every: int = 8):
    return LeniaRollout.apply(A0, mu, sigma, weight, betas, dt, world, T, ckpt_every)
