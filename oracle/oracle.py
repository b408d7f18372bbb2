"""ctypes wrapper around the FP64 Lenia oracle (oracle/lenia_oracle.c).

TEST INFRASTRUCTURE ONLY.  The tests, the smoke entry point and the CPU
baseline of ``bench.py`` use it as the checker; the product library never
loads it.

World and kernel descriptions are plain dicts with the same fields as the
library's ``lenia_world_desc`` (see include/lenia_mi355x.h), so one dict drives
both the oracle and the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblenia_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_random.so")

SHAPE_SHELL, SHAPE_RING = 0, 1
CORE_GAUSSIAN, CORE_EXP = 0, 1
TAG_PATCH = 1009


class LoKernel(C.Structure):
    _fields_ = [
        ("shape", C.c_int32),
        ("core", C.c_int32),
        ("radius", C.c_double),
        ("shellMu", C.c_double),
        ("shellSigma", C.c_double),
        ("nBetas", C.c_int32),
        ("betas", C.c_double * 4),
        ("mu", C.c_double),
        ("sigma", C.c_double),
        ("weight", C.c_double),
    ]


class LoWorld(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("growthMode", C.c_int32),
        ("dt", C.c_double),
        ("noiseAmount", C.c_double),
        ("muJitter", C.c_double),
        ("injectAmount", C.c_double),
        ("injectCount", C.c_int32),
        ("radius", C.c_int32),
        ("nKernels", C.c_int32),
        ("k", LoKernel * 4),
    ]


def build() -> None:
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-B", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32, f64p = C.c_uint32, C.POINTER(C.c_double)
        L.lo_pcg_hash.restype = u32
        L.lo_pcg_hash.argtypes = [u32]
        L.lo_derive_seed.restype = u32
        L.lo_derive_seed.argtypes = [u32, u32]
        L.lo_hash3.restype = u32
        L.lo_hash3.argtypes = [u32, u32, u32, u32]
        L.lo_rand01.restype = C.c_double
        L.lo_rand01.argtypes = [u32, u32, u32, u32]
        L.lo_module_seed.restype = u32
        L.lo_module_seed.argtypes = [u32]
        L.lo_step_seed.restype = u32
        L.lo_step_seed.argtypes = [u32, u32, u32]
        L.lo_splitmix_init.argtypes = [C.POINTER(C.c_uint64), C.c_uint64]
        L.lo_splitmix_next.restype = C.c_uint64
        L.lo_splitmix_next.argtypes = [C.POINTER(C.c_uint64)]
        L.lo_kernel_half.restype = C.c_int
        L.lo_kernel_half.argtypes = [C.POINTER(LoKernel)]
        L.lo_kernel_raw.restype = C.c_double
        L.lo_kernel_raw.argtypes = [C.POINTER(LoKernel), C.c_int, C.c_int]
        L.lo_kernel_window.restype = C.c_int
        L.lo_kernel_window.argtypes = [C.POINTER(LoKernel), C.c_int, f64p]
        L.lo_kernel_image.argtypes = [C.POINTER(LoKernel), C.c_int, C.c_int, f64p]
        L.lo_init_patch.argtypes = [C.c_int, C.c_int, u32, C.c_int, C.c_int, C.c_int, C.c_int, f64p]
        L.lo_init_value_noise.argtypes = [C.c_int, C.c_int, u32, C.c_double, C.c_double, f64p]
        L.lo_potential_direct.argtypes = [C.c_int, C.c_int, f64p, f64p, C.c_int, f64p]
        L.lo_fft.argtypes = [C.c_int, f64p, f64p, C.c_int]
        L.lo_fft2.argtypes = [C.c_int, C.c_int, f64p, f64p, C.c_int]
        L.lo_potential_fft.argtypes = [C.c_int, C.c_int, f64p, f64p, f64p]
        L.lo_step.restype = C.c_int
        L.lo_step.argtypes = [C.POINTER(LoWorld), f64p, f64p, C.c_int, u32, u32]
        L.lo_rollout.restype = C.c_int
        L.lo_rollout.argtypes = [C.POINTER(LoWorld), f64p, C.c_int, C.c_int, f64p]
        L.lo_rollout_grad.restype = C.c_int
        L.lo_rollout_grad.argtypes = [C.POINTER(LoWorld), f64p, C.c_int, f64p, f64p,
                                      f64p, f64p, f64p, f64p, f64p]
        L.lo_rollout_piece.restype = C.c_int
        L.lo_rollout_piece.argtypes = [C.POINTER(LoWorld), f64p, C.c_int, C.POINTER(C.c_uint8),
                                       C.c_int, f64p]
        L.lo_init_stamp.argtypes = [C.c_int, C.c_int, f64p, C.c_int, C.c_int, u32, u32, u32, f64p]
        L.lo_colormap.argtypes = [f64p, C.c_int, C.c_int, f64p, f64p, f64p, f64p, C.c_double,
                                  C.c_double, C.c_double, C.c_int, f64p, C.c_int, C.c_int]
        L.lo_threads.argtypes = [C.c_int]
        L.lo_max_threads.restype = C.c_int
        _lib = L
    return _lib


def ref_lib():
    """The reference's Random.h compiled into oracle/_ref, or None."""
    if not os.path.exists(REF_PATH):
        return None
    L = C.CDLL(REF_PATH)
    L.ref_pcg_hash.restype = C.c_uint32
    L.ref_pcg_hash.argtypes = [C.c_uint32]
    L.ref_derive_seed.restype = C.c_uint32
    L.ref_derive_seed.argtypes = [C.c_uint32, C.c_uint32]
    L.ref_splitmix_stream.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_uint64)]
    L.ref_splitmix_float01.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_float)]
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _c64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- descriptors
def kernel_struct(k: dict) -> LoKernel:
    s = LoKernel()
    s.shape = int(k.get("shape", SHAPE_RING))
    s.core = int(k.get("core", CORE_GAUSSIAN))
    s.radius = float(k["radius"])
    s.shellMu = float(k.get("shellMu", 0.5))
    s.shellSigma = float(k.get("shellSigma", 0.15))
    betas = list(k.get("betas", [1.0]))
    s.nBetas = len(betas)
    for i, b in enumerate(betas[:4]):
        s.betas[i] = float(b)
    s.mu = float(k.get("mu", 0.15))
    s.sigma = float(k.get("sigma", 0.017))
    s.weight = float(k.get("weight", 1.0))
    return s


def world_struct(w: dict) -> LoWorld:
    s = LoWorld()
    s.width = int(w["width"])
    s.height = int(w["height"])
    s.growthMode = int(w.get("growthMode", 0))
    s.dt = float(w.get("dt", 0.1))
    s.noiseAmount = float(w.get("noiseAmount", 0.0))
    s.muJitter = float(w.get("muJitter", 0.0))
    s.injectAmount = float(w.get("injectAmount", 0.0))
    s.injectCount = int(w.get("injectCount", 3))
    s.radius = int(w.get("radius", 13))
    ks = w["kernels"]
    s.nKernels = len(ks)
    for i, k in enumerate(ks[:4]):
        s.k[i] = kernel_struct(k)
    return s


# ---------------------------------------------------------------- functions
def pcg_hash(v: int) -> int:
    return lib().lo_pcg_hash(v & 0xFFFFFFFF)


def derive_seed(g: int, tag: int) -> int:
    return lib().lo_derive_seed(g & 0xFFFFFFFF, tag & 0xFFFFFFFF)


def module_seed(scene_seed: int) -> int:
    return lib().lo_module_seed(scene_seed & 0xFFFFFFFF)


def step_seed(module_seed_: int, frame: int, sub: int) -> int:
    return lib().lo_step_seed(module_seed_, frame & 0xFFFFFFFF, sub)


def rand01(x: int, y: int, tag: int, seed: int) -> float:
    return lib().lo_rand01(x, y, tag, seed)


def splitmix_stream(seed: int, n: int) -> list[int]:
    st = C.c_uint64()
    lib().lo_splitmix_init(C.byref(st), seed)
    return [lib().lo_splitmix_next(C.byref(st)) for _ in range(n)]


def kernel_half(k: dict) -> int:
    s = kernel_struct(k)
    return lib().lo_kernel_half(C.byref(s))


def kernel_window(k: dict, half: int | None = None) -> np.ndarray:
    s = kernel_struct(k)
    h = lib().lo_kernel_half(C.byref(s)) if half is None else half
    out = np.zeros((2 * h + 1, 2 * h + 1))
    lib().lo_kernel_window(C.byref(s), h, _p(out))
    return out


def kernel_image(k: dict, W: int, H: int) -> np.ndarray:
    s = kernel_struct(k)
    out = np.zeros((H, W))
    lib().lo_kernel_image(C.byref(s), W, H, _p(out))
    return out


def init_patch(W, H, seed, x0, y0, pw, ph) -> np.ndarray:
    out = np.zeros((H, W))
    lib().lo_init_patch(W, H, seed, x0, y0, pw, ph, _p(out))
    return out


def init_value_noise(W, H, seed, coverage=0.4, scale=24.0) -> np.ndarray:
    out = np.zeros((H, W))
    lib().lo_init_value_noise(W, H, seed, coverage, scale, _p(out))
    return out


def init_stamp(W, H, org, count, rotate, seed) -> np.ndarray:
    """leniaStampInit (Shaders/Field/Lenia.metal:242-303); org is rows x cols"""
    o = _c64(org)
    out = np.zeros((H, W))
    lib().lo_init_stamp(W, H, _p(o), o.shape[1], o.shape[0], count, 1 if rotate else 0, seed, _p(out))
    return out


def colormap(field, params: dict, dst_w: int | None = None, dst_h: int | None = None) -> np.ndarray:
    """colorMapField / colorMapFieldScaled (Shaders/Render/ColorMap.metal:17-75) in FP64:
    (dst_h, dst_w, 4) rgba"""
    f = _c64(field)
    H, W = f.shape
    dw, dh = dst_w or W, dst_h or H
    out = np.zeros((dh, dw, 4))
    vec = {k: (C.c_double * 3)(*[float(x) for x in params[k]]) for k in "abcd"}
    lib().lo_colormap(_p(f), W, H, vec["a"], vec["b"], vec["c"], vec["d"],
                      float(params.get("inputScale", 1.0)), float(params.get("inputBias", 0.0)),
                      float(params.get("exposure", 1.0)), int(params.get("channel", 0)), _p(out), dw, dh)
    return out


def luma_stats(rgba) -> tuple[float, float]:
    """computeLumaStats (Apps/LifePresetLab/main.cpp:67-86) of an RGBA layer"""
    px = np.asarray(rgba, np.float64).reshape(-1, 4)
    lum = 0.2126 * px[:, 0] + 0.7152 * px[:, 1] + 0.0722 * px[:, 2]
    mean = float(lum.sum() / lum.size)
    var = float((lum * lum).sum() / lum.size - mean * mean)
    return mean, max(var, 0.0) ** 0.5


def potential_direct(A, win) -> np.ndarray:
    A = _c64(A)
    win = _c64(win)
    H, W = A.shape
    half = (win.shape[0] - 1) // 2
    u = np.zeros_like(A)
    lib().lo_potential_direct(W, H, _p(A), _p(win), half, _p(u))
    return u


def potential_fft(A, img) -> np.ndarray:
    A = _c64(A)
    img = _c64(img)
    H, W = A.shape
    u = np.zeros_like(A)
    lib().lo_potential_fft(W, H, _p(A), _p(img), _p(u))
    return u


def fft(x: np.ndarray, inverse=False) -> np.ndarray:
    re = _c64(np.real(x)).copy()
    im = _c64(np.imag(x)).copy()
    lib().lo_fft(len(re), _p(re), _p(im), int(inverse))
    return re + 1j * im


def fft2(x: np.ndarray, inverse=False) -> np.ndarray:
    re = _c64(np.real(x)).copy()
    im = _c64(np.imag(x)).copy()
    H, W = re.shape
    lib().lo_fft2(W, H, _p(re), _p(im), int(inverse))
    return re + 1j * im


def step(world: dict, A, use_fft=False, seed=0, frame=0) -> np.ndarray:
    s = world_struct(world)
    A = _c64(A)
    out = np.zeros_like(A)
    rc = lib().lo_step(C.byref(s), _p(A), _p(out), int(use_fft), seed & 0xFFFFFFFF, frame)
    if rc != 0:
        raise ValueError("oracle step rejected the world description")
    return out


def rollout(world: dict, A0, T: int, use_fft=True) -> np.ndarray:
    s = world_struct(world)
    A0 = _c64(A0)
    out = np.zeros_like(A0)
    rc = lib().lo_rollout(C.byref(s), _p(A0), T, int(use_fft), _p(out))
    if rc != 0:
        raise ValueError("oracle rollout rejected the world description")
    return out


def rollout_piece(world: dict, A0, T: int, mask=None):
    """(A_T, mask): with mask None, the FFT-form rollout recording each step's
    clip state; given a recorded mask, the rollout with those clip states
    imposed (the smooth piece of the map the recorded trajectory lies on)"""
    s = world_struct(world)
    A0 = _c64(A0)
    out = np.zeros_like(A0)
    record = mask is None
    m = np.zeros((T,) + A0.shape, np.uint8) if record else np.ascontiguousarray(mask, np.uint8)
    if m.shape != (T,) + A0.shape:
        raise ValueError("mask shape does not match (T, H, W)")
    rc = lib().lo_rollout_piece(C.byref(s), _p(A0), T, m.ctypes.data_as(C.POINTER(C.c_uint8)),
                                int(record), _p(out))
    if rc != 0:
        raise ValueError("oracle rollout rejected the world description")
    return out, m


def rollout_grad(world: dict, A0, T: int, gT) -> dict:
    s = world_struct(world)
    A0 = _c64(A0)
    gT = _c64(gT)
    gA = np.zeros_like(A0)
    gmu, gsg, gw = np.zeros(4), np.zeros(4), np.zeros(4)
    gb = np.zeros(16)
    gdt = np.zeros(1)
    rc = lib().lo_rollout_grad(C.byref(s), _p(A0), T, _p(gT), _p(gA), _p(gmu), _p(gsg),
                               _p(gw), _p(gb), _p(gdt))
    if rc != 0:
        raise ValueError("oracle gradient rejected the world description")
    K = s.nKernels
    return {"A0": gA, "mu": gmu[:K], "sigma": gsg[:K], "weight": gw[:K],
            "betas": gb.reshape(4, 4)[:K], "dt": float(gdt[0])}


def set_threads(n: int) -> None:
    lib().lo_threads(n)


def max_threads() -> int:
    return lib().lo_max_threads()
