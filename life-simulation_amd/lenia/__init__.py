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
                   CORE_GAUSSIAN, INIT_NOISE, INIT_NONE, INIT_PATCH, INIT_STAMP, SHAPE_RING,
                   SHAPE_SHELL, LeniaError, check)

__all__ = ["Context", "World", "Batch", "Scene", "Sweep", "LeniaError", "describe_scene",
           "plan_autotune", "plan_cache_set", "plan_cache_get", "plan_cache_clear",
           "plan_cache_save", "plan_cache_load",
           "colormap_defaults", "colormap_field", "CONV_DIRECT", "CONV_FFT", "CONV_DIRECT_BF16",
           "CONV_DIRECT_BF16X3", "SHAPE_SHELL", "SHAPE_RING", "CORE_GAUSSIAN", "CORE_EXP",
           "INIT_PATCH", "INIT_NOISE", "INIT_NONE", "INIT_STAMP"]


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

    @classmethod
    def _borrow(cls, ctx: Context, handle: int) -> "World":
        """a non-owning view of a world owned by a scene or sweep"""
        w = cls.__new__(cls)
        w.ctx = ctx
        w._h = C.c_void_p(handle)
        w._rollouts = 0
        w._owned = False
        d = _abi.WorldDesc()
        check(_abi.lib().lenia_world_get_desc(w._h, C.byref(d)))
        w.desc = _abi.desc_dict(d)
        w.W, w.H = int(d.width), int(d.height)
        return w

    @property
    def handle(self):
        return self._h

    def set_organism(self, cells, stamp_count: int = 6, stamp_rotate: bool = True):
        """organism cells (rows x cols in [0,1]) for INIT_STAMP"""
        a = np.ascontiguousarray(cells, dtype=np.float32)
        check(_abi.lib().lenia_world_set_organism(self._h, _fptr(a), a.shape[1], a.shape[0],
                                                  stamp_count, 1 if stamp_rotate else 0))

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

    def save_npy(self, path: str) -> None:
        check(_abi.lib().lenia_world_save_npy(self._h, str(path).encode()))

    def load_npy(self, path: str) -> None:
        check(_abi.lib().lenia_world_load_npy(self._h, str(path).encode()))

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
                "rowThreads": d.rowThreads, "colThreads": d.colThreads,
                "rowEngine": d.rowEngine, "colEngine": d.colEngine}

    def close(self):
        if self._h and getattr(self, "_owned", True):
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

    def frames(self, frame0: int, n_frames: int, substeps: int = 1, graph_steps: int = 0):
        """n_frames consecutive frames of `substeps` steps each, from frame0"""
        check(_abi.lib().lenia_batch_frames(self._h, frame0, n_frames, substeps, graph_steps))

    def autotune(self, n_steps: int = 10) -> float:
        """time every launch shape valid for all worlds, keep the fastest;
        returns the tuned kernel ms per step (states and clocks restored)"""
        ms = C.c_double()
        check(_abi.lib().lenia_batch_autotune(self._h, n_steps, C.byref(ms)))
        return ms.value

    def set_grids(self, col_per_cu: int, row_per_cu: int):
        """resident workgroups per CU taking tiles from a queue, per pass
        (0: one workgroup per tile)"""
        check(_abi.lib().lenia_batch_set_grids(self._h, col_per_cu, row_per_cu))

    def grids(self) -> tuple[int, int]:
        """(column, row) workgroups per CU; 0 is one workgroup per tile"""
        c, r = C.c_uint32(), C.c_uint32()
        check(_abi.lib().lenia_batch_get_grids(self._h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def set_waves(self, row_waves: int, col_waves: int):
        """waves per workgroup of the wave executor per pass (0: every world on
        the workgroup-LDS executor); same bits either way"""
        check(_abi.lib().lenia_batch_set_waves(self._h, row_waves, col_waves))

    def waves(self) -> tuple[int, int]:
        r, c = C.c_uint32(), C.c_uint32()
        check(_abi.lib().lenia_batch_get_waves(self._h, C.byref(r), C.byref(c)))
        return r.value, c.value

    def close(self):
        if self._h:
            _abi.lib().lenia_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


ROWS, COLS = 0, 1  # plan-cache axes


def plan_autotune(ctx: Context, desc: dict, n_steps: int = 10):
    """time every compiled radix plan of desc's width and height; the fastest
    go into the plan cache (worlds created afterwards use them).  Returns
    (planW, planH, kernel ms per step)."""
    d = _abi.world_desc(desc)
    pw, ph, ms = C.c_uint32(), C.c_uint32(), C.c_double()
    check(_abi.lib().lenia_plan_autotune(ctx.handle, C.byref(d), n_steps, C.byref(pw), C.byref(ph),
                                         C.byref(ms)))
    return pw.value, ph.value, ms.value


def plan_cache_set(axis: int, n: int, plan: int) -> None:
    check(_abi.lib().lenia_plan_cache_set(axis, n, plan))


def plan_cache_get(axis: int, n: int) -> int:
    v = C.c_uint32()
    check(_abi.lib().lenia_plan_cache_get(axis, n, C.byref(v)))
    return v.value


def plan_cache_clear() -> None:
    _abi.lib().lenia_plan_cache_clear()


def plan_cache_save(path: str) -> None:
    check(_abi.lib().lenia_plan_cache_save(str(path).encode()))


def plan_cache_load(path: str) -> None:
    check(_abi.lib().lenia_plan_cache_load(str(path).encode()))


def describe_scene(text: str, module: int = 0):
    """(descriptor dict, colour-map dict, module name) that module `module`
    of a Lenia scene JSON resolves to; no GPU needed"""
    b = text.encode()
    d, c = _abi.WorldDesc(), _abi.ColorMap()
    name = C.create_string_buffer(256)
    check(_abi.lib().lenia_scene_describe(b, len(b), module, C.byref(d), C.byref(c), name, 256))
    return _abi.desc_dict(d), _abi.colormap_dict(c), name.value.decode()


def colormap_defaults() -> dict:
    c = _abi.ColorMap()
    _abi.lib().lenia_colormap_lenia_defaults(C.byref(c))
    return _abi.colormap_dict(c)


def colormap_field(ctx: Context, field_ptr: int, src_w: int, src_h: int, params: dict,
                   rgba_ptr: int, dst_w: int, dst_h: int) -> None:
    """device field (src_h x src_w float32) -> device RGBA16F (dst_h x dst_w x 4)"""
    c = _abi.colormap_struct(params)
    check(_abi.lib().lenia_colormap_field(ctx.handle, C.c_void_p(field_ptr), src_w, src_h,
                                          C.byref(c), C.c_void_p(rgba_ptr), dst_w, dst_h))


class Scene:
    """A Lenia scene JSON stepped frame by frame (SceneRunner for Lenia modules)."""

    def __init__(self, ctx: Context, text: str, substeps: int = 1):
        self.ctx = ctx
        b = text.encode()
        self._h = C.c_void_p()
        check(_abi.lib().lenia_scene_create(ctx.handle, b, len(b), substeps, C.byref(self._h)))

    def info(self) -> dict:
        i = _abi.SceneInfo()
        check(_abi.lib().lenia_scene_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _abi.SceneInfo._fields_}

    def reset(self):
        check(_abi.lib().lenia_scene_reset(self._h))

    def step(self, n_frames: int = 1):
        check(_abi.lib().lenia_scene_step(self._h, n_frames))

    def set_param(self, target: str, value: float):
        check(_abi.lib().lenia_scene_set_param(self._h, target.encode(), float(value)))

    def world(self, module: int = 0) -> World:
        h = C.c_void_p()
        check(_abi.lib().lenia_scene_world(self._h, module, C.byref(h)))
        return World._borrow(self.ctx, h.value)

    def layer(self, module: int = 0) -> np.ndarray:
        """the module's colour layer, (height, width, 4) float16"""
        i = self.info()
        out = np.empty((i["height"], i["width"], 4), np.float16)
        check(_abi.lib().lenia_scene_layer(self._h, module, out.ctypes.data_as(C.c_void_p),
                                           out.size, 0))
        return out

    def close(self):
        if self._h:
            _abi.lib().lenia_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Sweep:
    """Variants of a one-module Lenia scene (PresetLab sweep JSON) on the batched launch."""

    def __init__(self, ctx: Context, scene_text: str, sweep_text: str, substeps: int = 1):
        self.ctx = ctx
        a, b = scene_text.encode(), sweep_text.encode()
        self._h = C.c_void_p()
        check(_abi.lib().lenia_sweep_create(ctx.handle, a, len(a), b, len(b), substeps,
                                            C.byref(self._h)))
        n = C.c_uint32()
        check(_abi.lib().lenia_sweep_count(self._h, C.byref(n)))
        self.n = n.value

    def run(self, n_frames: int, graph_steps: int = 0):
        check(_abi.lib().lenia_sweep_run(self._h, n_frames, graph_steps))

    def activity(self):
        """(mean luma, stddev luma, alive) arrays of the last run"""
        m = (C.c_double * self.n)()
        sd = (C.c_double * self.n)()
        al = (C.c_int * self.n)()
        check(_abi.lib().lenia_sweep_activity(self._h, m, sd, al, self.n))
        return np.array(m[:]), np.array(sd[:]), np.array(al[:], bool)

    def world(self, v: int) -> World:
        h = C.c_void_p()
        check(_abi.lib().lenia_sweep_world(self._h, v, C.byref(h)))
        if not h.value:
            raise LeniaError(-1, f"variant {v} has no world (see its error in the report)")
        return World._borrow(self.ctx, h.value)

    def report(self) -> dict:
        import json
        need = C.c_size_t()
        check(_abi.lib().lenia_sweep_report(self._h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        check(_abi.lib().lenia_sweep_report(self._h, buf, need.value, C.byref(need)))
        return json.loads(buf.value.decode())

    def close(self):
        if self._h:
            _abi.lib().lenia_sweep_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
