"""Python handles over the MI355X Lenia C ABI (include/lenia_mi355x.h).

``Context`` / ``World`` / ``Batch`` wrap the opaque handles; arrays cross the
boundary as numpy float32.  The compute always runs in liblenia_mi355x.so;
there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import (CONV_DIRECT, CONV_DIRECT_BF16, CONV_DIRECT_BF16X3, CONV_FFT, CORE_EXP,
                   CORE_GAUSSIAN, INIT_NOISE, INIT_NONE, INIT_PATCH, SHAPE_RING, SHAPE_SHELL,
                   LeniaError, check)

__all__ = ["Context", "World", "Batch", "LeniaError", "CONV_DIRECT", "CONV_FFT",
           "CONV_DIRECT_BF16", "CONV_DIRECT_BF16X3", "SHAPE_SHELL", "SHAPE_RING",
           "CORE_GAUSSIAN", "CORE_EXP", "INIT_PATCH", "INIT_NOISE", "INIT_NONE"]


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Context:
    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(_abi.lib().lenia_ctx_create(device, C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def sync(self):
        check(_abi.lib().lenia_ctx_sync(self._h))

    def stream(self) -> int:
        s = C.c_void_p()
        check(_abi.lib().lenia_ctx_stream(self._h, C.byref(s)))
        return s.value or 0

    def memory(self):
        b, n = C.c_uint64(), C.c_uint32()
        check(_abi.lib().lenia_ctx_memory(self._h, C.byref(b), C.byref(n)))
        return b.value, n.value

    def last_gpu_ms(self) -> float:
        ms = C.c_double()
        check(_abi.lib().lenia_ctx_last_gpu_ms(self._h, C.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            _abi.lib().lenia_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class World:
    def __init__(self, ctx: Context, desc: dict):
        self.ctx = ctx
        self.desc = dict(desc)
        self.W, self.H = int(desc["width"]), int(desc["height"])
        self._h = C.c_void_p()
        self._rollouts = 0  # forward sweeps run (each replaces the kept checkpoints)
        d = _abi.world_desc(desc)
        check(_abi.lib().lenia_world_create(ctx.handle, C.byref(d), C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def reset(self, scene_seed: int):
        check(_abi.lib().lenia_world_reset(self._h, scene_seed & 0xFFFFFFFF))

    def frame(self, frame_index: int, substeps: int = 1):
        check(_abi.lib().lenia_world_frame(self._h, frame_index & 0xFFFFFFFF, substeps))

    def step(self, n: int = 1):
        check(_abi.lib().lenia_world_step(self._h, n))

    def set_params(self, desc: dict):
        d = _abi.world_desc(desc)
        check(_abi.lib().lenia_world_set_params(self._h, C.byref(d)))
        self.desc = dict(desc)

    def read(self) -> np.ndarray:
        out = np.empty((self.H, self.W), np.float32)
        check(_abi.lib().lenia_world_read_state(self._h, _fptr(out), out.size))
        return out

    def write(self, A) -> None:
        a = np.ascontiguousarray(A, dtype=np.float32)
        if a.shape != (self.H, self.W):
            raise ValueError(f"state shape {a.shape} != {(self.H, self.W)}")
        check(_abi.lib().lenia_world_write_state(self._h, _fptr(a), a.size))

    def potential(self, k: int = 0) -> np.ndarray:
        out = np.empty((self.H, self.W), np.float32)
        check(_abi.lib().lenia_world_potential(self._h, k, _fptr(out), out.size))
        return out

    def state_device_ptr(self) -> int:
        p = C.c_void_p()
        check(_abi.lib().lenia_world_state_device(self._h, C.byref(p)))
        return p.value or 0

    def pass_times(self, n: int) -> list[float]:
        """mean ms per pass over n steps ([column, row] for FFT, [step] direct)"""
        out = (C.c_double * 4)()
        npass = C.c_uint32()
        check(_abi.lib().lenia_world_pass_times(self._h, n, out, 4, C.byref(npass)))
        return [out[i] for i in range(npass.value)]

    def rollout(self, A0, T: int, ckpt_every: int = 0) -> np.ndarray:
        """T steps from A0 (host arrays); returns A_T and keeps the checkpoints
        for rollout_grad(None, ...)"""
        a = np.ascontiguousarray(A0, dtype=np.float32)
        out = np.empty_like(a)
        self._rollouts += 1
        check(_abi.lib().lenia_rollout_forward(self._h, T, ckpt_every, a.ctypes.data,
                                               out.ctypes.data, 0))
        return out

    def version_token(self) -> int:
        """changes whenever a forward sweep replaces the kept checkpoints"""
        return self._rollouts

    def rollout_grad(self, A0, T: int, gAT, ckpt_every: int = 0) -> dict:
        """reverse pass of rollout(A0, T) for dL/dA_T = gAT (host arrays);
        A0 None reuses the checkpoints of the last rollout()"""
        g = np.ascontiguousarray(gAT, dtype=np.float32)
        a = None if A0 is None else np.ascontiguousarray(A0, dtype=np.float32)
        gA0 = np.empty_like(g)
        pg = _abi.ParamGrads()
        if a is not None:
            self._rollouts += 1
        check(_abi.lib().lenia_rollout_backward(self._h, T, ckpt_every,
                                                None if a is None else a.ctypes.data,
                                                g.ctypes.data, gA0.ctypes.data, C.byref(pg), 0))
        return grads_dict(pg, len(self.desc["kernels"]), gA0)

    def autotune(self, n_steps: int = 20) -> float:
        """time every launch shape the FFT plans allow, keep the fastest;
        returns the tuned kernel ms per step (state and clock restored)"""
        ms = C.c_double()
        check(_abi.lib().lenia_world_autotune(self._h, n_steps, C.byref(ms)))
        return ms.value

    def resolved(self) -> dict:
        d = _abi.WorldDesc()
        check(_abi.lib().lenia_world_get_desc(self._h, C.byref(d)))
        return {"fftPlanW": d.fftPlanW, "fftPlanH": d.fftPlanH,
                "rowPairsPerTile": d.rowPairsPerTile, "colsPerTile": d.colsPerTile,
                "rowThreads": d.rowThreads, "colThreads": d.colThreads}

    def close(self):
        if self._h:
            _abi.lib().lenia_world_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def grads_dict(pg, K: int, gA0=None) -> dict:
    out = {"mu": np.array(pg.dMu[:K]), "sigma": np.array(pg.dSigma[:K]),
           "weight": np.array(pg.dWeight[:K]),
           "betas": np.array([[pg.dBetas[k][j] for j in range(4)] for k in range(K)]),
           "dt": pg.dDt}
    if gA0 is not None:
        out["A0"] = gA0
    return out


class Batch:
    def __init__(self, ctx: Context, worlds: list[World]):
        self.ctx = ctx
        self.worlds = list(worlds)
        arr = (C.c_void_p * len(worlds))(*[w.handle.value for w in worlds])
        self._h = C.c_void_p()
        check(_abi.lib().lenia_batch_create(ctx.handle, arr, len(worlds), C.byref(self._h)))

    def step(self, n: int, graph_steps: int = 0):
        check(_abi.lib().lenia_batch_step(self._h, n, graph_steps))

    def autotune(self, n_steps: int = 10) -> float:
        """time every launch shape valid for all worlds, keep the fastest;
        returns the tuned kernel ms per step (states and clocks restored)"""
        ms = C.c_double()
        check(_abi.lib().lenia_batch_autotune(self._h, n_steps, C.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            _abi.lib().lenia_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
