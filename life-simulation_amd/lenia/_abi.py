"""ctypes binding of liblenia_mi355x.so (the C ABI in include/lenia_mi355x.h).

This module only loads the library and mirrors its structs; it never falls
back to anything else.  If the shared library is missing, every entry point
raises.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "liblenia_mi355x.so")

OK = 0
ERR_INVALID_ARG = -1
ERR_UNSUPPORTED = -2
ERR_DEVICE = -3
ERR_OUT_OF_MEMORY = -4
ERR_STATE = -5

CONV_DIRECT, CONV_FFT, CONV_DIRECT_BF16, CONV_DIRECT_BF16X3 = 0, 1, 2, 3
SHAPE_SHELL, SHAPE_RING = 0, 1
CORE_GAUSSIAN, CORE_EXP = 0, 1
INIT_PATCH, INIT_NOISE, INIT_NONE, INIT_STAMP = 0, 1, 2, 3

# every symbol include/lenia_mi355x.h declares
EXPORTS = [
    "lenia_last_error", "lenia_version",
    "lenia_ctx_create", "lenia_ctx_destroy", "lenia_ctx_sync", "lenia_ctx_stream",
    "lenia_ctx_memory", "lenia_ctx_last_gpu_ms",
    "lenia_world_create", "lenia_world_destroy", "lenia_world_reset", "lenia_world_frame",
    "lenia_world_step", "lenia_world_set_params", "lenia_world_read_state",
    "lenia_world_write_state", "lenia_world_state_device", "lenia_world_potential",
    "lenia_world_get_desc", "lenia_world_pass_times",
    "lenia_batch_create", "lenia_batch_destroy", "lenia_batch_step",
    "lenia_batch_set_grids", "lenia_batch_get_grids", "lenia_batch_set_waves", "lenia_batch_get_waves",
    "lenia_rollout_forward", "lenia_rollout_backward",
    "lenia_world_autotune", "lenia_batch_autotune", "lenia_world_set_organism",
    "lenia_world_save_npy", "lenia_world_load_npy",
    "lenia_batch_frames", "lenia_colormap_lenia_defaults", "lenia_colormap_field",
    "lenia_scene_describe", "lenia_scene_create", "lenia_scene_destroy", "lenia_scene_reset",
    "lenia_scene_step", "lenia_scene_set_param", "lenia_scene_get_info", "lenia_scene_world",
    "lenia_scene_layer", "lenia_sweep_create", "lenia_sweep_destroy", "lenia_sweep_count",
    "lenia_sweep_run", "lenia_sweep_activity", "lenia_sweep_world", "lenia_sweep_report",
    "lenia_fft_step_scratch", "lenia_plan_autotune", "lenia_plan_cache_set",
    "lenia_plan_cache_get", "lenia_plan_cache_clear", "lenia_plan_cache_save",
    "lenia_plan_cache_load",
]


class KernelDesc(C.Structure):
    _fields_ = [
        ("shape", C.c_uint32),
        ("core", C.c_uint32),
        ("radius", C.c_float),
        ("shellMu", C.c_float),
        ("shellSigma", C.c_float),
        ("nBetas", C.c_uint32),
        ("betas", C.c_float * 4),
        ("mu", C.c_float),
        ("sigma", C.c_float),
        ("weight", C.c_float),
    ]


class WorldDesc(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("convMode", C.c_uint32),
        ("growthMode", C.c_uint32),
        ("dt", C.c_float),
        ("noiseAmount", C.c_float),
        ("muJitter", C.c_float),
        ("injectAmount", C.c_float),
        ("injectCount", C.c_uint32),
        ("radius", C.c_uint32),
        ("nKernels", C.c_uint32),
        ("kernels", KernelDesc * 4),
        ("initMode", C.c_uint32),
        ("initX0", C.c_uint32),
        ("initY0", C.c_uint32),
        ("initW", C.c_uint32),
        ("initH", C.c_uint32),
        ("initCoverage", C.c_float),
        ("initScale", C.c_float),
        ("fftPlanW", C.c_uint32),
        ("fftPlanH", C.c_uint32),
        ("rowPairsPerTile", C.c_uint32),
        ("colsPerTile", C.c_uint32),
        ("rowThreads", C.c_uint32),
        ("colThreads", C.c_uint32),
        ("rowEngine", C.c_uint32),
        ("colEngine", C.c_uint32),
    ]


class ColorMap(C.Structure):
    _fields_ = [
        ("a", C.c_float * 3),
        ("b", C.c_float * 3),
        ("c", C.c_float * 3),
        ("d", C.c_float * 3),
        ("inputScale", C.c_float),
        ("inputBias", C.c_float),
        ("exposure", C.c_float),
        ("channel", C.c_uint32),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("seed", C.c_uint32),
        ("substeps", C.c_uint32),
        ("frameIndex", C.c_uint32),
        ("nModules", C.c_uint32),
    ]


class ParamGrads(C.Structure):
    _fields_ = [
        ("dMu", C.c_double * 4),
        ("dSigma", C.c_double * 4),
        ("dWeight", C.c_double * 4),
        ("dBetas", (C.c_double * 4) * 4),
        ("dDt", C.c_double),
    ]


class LeniaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lenia error {code}: {msg}")
        self.code = code


_lib = None
# True if torch was imported when the library was loaded.  torch bundles its
# own libamdhip64 (soname libamdhip64.so.7): loaded after torch, the library
# binds to that runtime; loaded before, it binds to /opt/rocm's, and a later
# torch import would bring a second HIP runtime into the process, whose device
# pointers the library cannot use.
loaded_after_torch = False


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib, loaded_after_torch
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    loaded_after_torch = "torch" in sys.modules
    L = C.CDLL(LIB_PATH)
    vp, i32, u32, sz = C.c_void_p, C.c_int, C.c_uint32, C.c_size_t
    fp = C.POINTER(C.c_float)
    sig = {
        "lenia_last_error": (C.c_char_p, []),
        "lenia_version": (C.c_char_p, []),
        "lenia_ctx_create": (i32, [i32, C.POINTER(vp)]),
        "lenia_ctx_destroy": (i32, [vp]),
        "lenia_ctx_sync": (i32, [vp]),
        "lenia_ctx_stream": (i32, [vp, C.POINTER(vp)]),
        "lenia_ctx_memory": (i32, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
        "lenia_ctx_last_gpu_ms": (i32, [vp, C.POINTER(C.c_double)]),
        "lenia_world_create": (i32, [vp, C.POINTER(WorldDesc), C.POINTER(vp)]),
        "lenia_world_destroy": (i32, [vp]),
        "lenia_world_reset": (i32, [vp, u32]),
        "lenia_world_frame": (i32, [vp, u32, u32]),
        "lenia_world_step": (i32, [vp, u32]),
        "lenia_world_set_params": (i32, [vp, C.POINTER(WorldDesc)]),
        "lenia_world_read_state": (i32, [vp, fp, sz]),
        "lenia_world_write_state": (i32, [vp, fp, sz]),
        "lenia_world_state_device": (i32, [vp, C.POINTER(vp)]),
        "lenia_world_potential": (i32, [vp, u32, fp, sz]),
        "lenia_world_get_desc": (i32, [vp, C.POINTER(WorldDesc)]),
        "lenia_world_pass_times": (i32, [vp, u32, C.POINTER(C.c_double), u32,
                                         C.POINTER(C.c_uint32)]),
        "lenia_batch_create": (i32, [vp, C.POINTER(vp), u32, C.POINTER(vp)]),
        "lenia_batch_destroy": (i32, [vp]),
        "lenia_batch_step": (i32, [vp, u32, u32]),
        "lenia_rollout_forward": (i32, [vp, u32, u32, vp, vp, i32]),
        "lenia_world_autotune": (i32, [vp, u32, C.POINTER(C.c_double)]),
        "lenia_batch_autotune": (i32, [vp, u32, C.POINTER(C.c_double)]),
        "lenia_batch_set_grids": (i32, [vp, u32, u32]),
        "lenia_batch_get_grids": (i32, [vp, C.POINTER(u32), C.POINTER(u32)]),
        "lenia_batch_set_waves": (i32, [vp, u32, u32]),
        "lenia_batch_get_waves": (i32, [vp, C.POINTER(u32), C.POINTER(u32)]),
        "lenia_rollout_backward": (i32, [vp, u32, u32, vp, vp, vp, C.POINTER(ParamGrads), i32]),
        "lenia_world_set_organism": (i32, [vp, fp, u32, u32, u32, u32]),
        "lenia_world_save_npy": (i32, [vp, C.c_char_p]),
        "lenia_world_load_npy": (i32, [vp, C.c_char_p]),
        "lenia_batch_frames": (i32, [vp, u32, u32, u32, u32]),
        "lenia_colormap_lenia_defaults": (None, [C.POINTER(ColorMap)]),
        "lenia_colormap_field": (i32, [vp, vp, u32, u32, C.POINTER(ColorMap), vp, u32, u32]),
        "lenia_scene_describe": (i32, [C.c_char_p, sz, u32, C.POINTER(WorldDesc), C.POINTER(ColorMap),
                                       C.c_char_p, sz]),
        "lenia_scene_create": (i32, [vp, C.c_char_p, sz, u32, C.POINTER(vp)]),
        "lenia_scene_destroy": (i32, [vp]),
        "lenia_scene_reset": (i32, [vp]),
        "lenia_scene_step": (i32, [vp, u32]),
        "lenia_scene_set_param": (i32, [vp, C.c_char_p, C.c_float]),
        "lenia_scene_get_info": (i32, [vp, C.POINTER(SceneInfo)]),
        "lenia_scene_world": (i32, [vp, u32, C.POINTER(vp)]),
        "lenia_scene_layer": (i32, [vp, u32, vp, sz, i32]),
        "lenia_sweep_create": (i32, [vp, C.c_char_p, sz, C.c_char_p, sz, u32, C.POINTER(vp)]),
        "lenia_sweep_destroy": (i32, [vp]),
        "lenia_sweep_count": (i32, [vp, C.POINTER(C.c_uint32)]),
        "lenia_sweep_run": (i32, [vp, u32, u32]),
        "lenia_sweep_activity": (i32, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int), u32]),
        "lenia_sweep_world": (i32, [vp, u32, C.POINTER(vp)]),
        "lenia_sweep_report": (i32, [vp, C.c_char_p, sz, C.POINTER(sz)]),
        "lenia_fft_step_scratch": (i32, [vp, C.POINTER(WorldDesc), u32, u32, vp, u32]),
        "lenia_plan_autotune": (i32, [vp, C.POINTER(WorldDesc), u32, C.POINTER(u32), C.POINTER(u32),
                                      C.POINTER(C.c_double)]),
        "lenia_plan_cache_set": (i32, [u32, u32, u32]),
        "lenia_plan_cache_get": (i32, [u32, u32, C.POINTER(u32)]),
        "lenia_plan_cache_clear": (None, []),
        "lenia_plan_cache_save": (i32, [C.c_char_p]),
        "lenia_plan_cache_load": (i32, [C.c_char_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != OK:
        raise LeniaError(rc, lib().lenia_last_error().decode())


def kernel_desc(k: dict) -> KernelDesc:
    d = KernelDesc()
    d.shape = int(k.get("shape", SHAPE_RING))
    d.core = int(k.get("core", CORE_GAUSSIAN))
    d.radius = float(k["radius"])
    d.shellMu = float(k.get("shellMu", 0.5))
    d.shellSigma = float(k.get("shellSigma", 0.15))
    betas = list(k.get("betas", [1.0]))
    d.nBetas = len(betas)
    for i, b in enumerate(betas[:4]):
        d.betas[i] = float(b)
    d.mu = float(k.get("mu", 0.15))
    d.sigma = float(k.get("sigma", 0.017))
    d.weight = float(k.get("weight", 1.0))
    return d


def desc_dict(d: WorldDesc) -> dict:
    """The dict form of a WorldDesc (inverse of world_desc)."""
    ks = []
    for i in range(d.nKernels):
        k = d.kernels[i]
        ks.append({"shape": k.shape, "core": k.core, "radius": k.radius, "shellMu": k.shellMu,
                   "shellSigma": k.shellSigma, "betas": [k.betas[j] for j in range(k.nBetas)],
                   "mu": k.mu, "sigma": k.sigma, "weight": k.weight})
    out = {f: getattr(d, f) for f, _ in WorldDesc._fields_ if f != "kernels"}
    out["kernels"] = ks
    return out


def colormap_dict(c: ColorMap) -> dict:
    return {"a": list(c.a), "b": list(c.b), "c": list(c.c), "d": list(c.d),
            "inputScale": c.inputScale, "inputBias": c.inputBias, "exposure": c.exposure,
            "channel": c.channel}


def colormap_struct(p: dict) -> ColorMap:
    c = ColorMap()
    lib().lenia_colormap_lenia_defaults(C.byref(c))
    for key in ("a", "b", "c", "d"):
        if key in p:
            for i in range(3):
                getattr(c, key)[i] = float(p[key][i])
    for key in ("inputScale", "inputBias", "exposure"):
        if key in p:
            setattr(c, key, float(p[key]))
    if "channel" in p:
        c.channel = int(p["channel"])
    return c


def world_desc(w: dict) -> WorldDesc:
    """Builds a WorldDesc from the dict form shared with the oracle."""
    d = WorldDesc()
    d.width = int(w["width"])
    d.height = int(w["height"])
    d.convMode = int(w.get("convMode", CONV_DIRECT))
    d.growthMode = int(w.get("growthMode", 0))
    d.dt = float(w.get("dt", 0.1))
    d.noiseAmount = float(w.get("noiseAmount", 0.0))
    d.muJitter = float(w.get("muJitter", 0.0))
    d.injectAmount = float(w.get("injectAmount", 0.0))
    d.injectCount = int(w.get("injectCount", 3))
    d.radius = int(w.get("radius", 13))
    ks = w["kernels"]
    d.nKernels = len(ks)
    for i, k in enumerate(ks[:4]):
        d.kernels[i] = kernel_desc(k)
    d.initMode = int(w.get("initMode", INIT_PATCH))
    x0, y0, pw, ph = w.get("initRect", (0, 0, 0, 0))
    d.initX0, d.initY0, d.initW, d.initH = int(x0), int(y0), int(pw), int(ph)
    d.initCoverage = float(w.get("initCoverage", 0.4))
    d.initScale = float(w.get("initScale", 24.0))
    d.fftPlanW = int(w.get("fftPlanW", 0))
    d.fftPlanH = int(w.get("fftPlanH", 0))
    d.rowPairsPerTile = int(w.get("rowPairsPerTile", 0))
    d.colsPerTile = int(w.get("colsPerTile", 0))
    d.rowThreads = int(w.get("rowThreads", 0))
    d.colThreads = int(w.get("colThreads", 0))
    d.rowEngine = int(w.get("rowEngine", 0))
    d.colEngine = int(w.get("colEngine", 0))
    return d
