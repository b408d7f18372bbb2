"""Lenia field-update benchmark on one MI355X (the LifeBench measurement,
Apps/LifeBench/main.cpp:43-177, for the Lenia hot path).

Workload (the headline metric's configuration): FFT-form Lenia, 1024x512 FP32
toroidal field, three ring kernels R=24 with shell peaks [1], [1,.5],
[.5,1,.25], weights .5/.3/.2, growth (mu, sigma) (.15,.015) (.26,.036)
(.12,.012), dt .1, seeded 256x256 patch init.  A step is one full field
update; the field stays resident in HBM.

Prints one JSON line:  value = Mcell-steps/s = W*H*steps / GPU seconds.
  roofline     dominant kernel (by mean launch time over a fixed sample of
               PASS_SAMPLES steps, HIP events around each kernel on the
               library's stream): algorithmic bytes per launch / that time;
               both passes are reported
  live_field   the same kernels on a field that stays alive (the headline
               world is empty after 18 steps; the FFT step's work does not
               depend on the values)
  batched      256 uneven worlds in one batch vs one world per launch
  matrix       the other configurations: 512^2 direct R=13, 1920x1080 and
               4096^2 FFT, each with its roofline and a CPU sample
  lifebench    per-frame GPU time of the reference's LifeBench scenes
  cpu_baseline the FP64 oracle's FFT step on the host cores, bounded sample
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "life-simulation_amd"), os.path.join(ROOT, "tests"),
          os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak, published
VALU_FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector peak, published
PASS_SAMPLES = 1000  # steps of per-kernel event timing behind every roofline figure
SCENES = os.path.join(ROOT, "tests", "golden", "scenes")


def fft_bytes_per_cell(W: int, K: int) -> dict:
    """Algorithmic HBM bytes per cell-step of each FFT pass (fp32 state,
    complex64 spectra, real fp32 kernel spectra, h = (W/2+1)/W):
      column pass: read F (8h) + read K spectra (4hK) + write K spectra (8hK)
      row pass:    read K spectra (8hK) + read/write state (8) + write F (8h)
    """
    h = (W // 2 + 1) / W
    cols = 8 * h + 12 * h * K
    rows = 8 * h * K + 8 + 8 * h
    return {"cols": cols, "rows": rows, "step": cols + rows}


def window_nnz(kernel: dict) -> int:
    """taps of a kernel window that are not zero (the algorithmic flops of the
    direct forms are 2 x nnz per cell-step)"""
    import numpy as np

    import oracle as O

    return int(np.count_nonzero(O.kernel_window(kernel)))


ENGINE_LDS, ENGINE_WAVE, ENGINE_WAVE_SPLIT = 1, 2, 3
ENGINE_NAMES = {ENGINE_LDS: "lds", ENGINE_WAVE: "wave", ENGINE_WAVE_SPLIT: "wave-split"}


def pass_kernel(name: str, engine: int) -> str:
    """kernel that runs a step pass on its executor: workgroup-LDS, wave, or
    kernel-split wave"""
    if engine == ENGINE_WAVE_SPLIT:
        return f"k_wave_{name}_split"
    return ("k_wave_" if engine == ENGINE_WAVE else "k_fft_") + name


def fft_roofline(W: int, H: int, K: int, passes: list, engines=(ENGINE_LDS, ENGINE_LDS)) -> dict:
    """per-pass achieved HBM bandwidth from the mean kernel durations;
    engines: the (column, row) pass executors"""
    bpc = fft_bytes_per_cell(W, K)
    out = {}
    for name, ms, eng in zip(("cols", "rows"), passes, engines):
        alg = bpc[name] * W * H
        gbs = alg / (ms * 1e-3) / 1e9
        out[name] = {"kernel": pass_kernel(name, eng), "ms": round(ms, 5),
                     "algorithmic_bytes_per_launch": int(round(alg)),
                     "algorithmic_bytes_per_cell": round(bpc[name], 3),
                     "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def cpu_sample(desc: dict, A, use_fft: bool, budget_s: float) -> dict:
    """FP64 oracle steps on all host threads for about budget_s seconds"""
    import oracle as O

    threads = O.max_threads()
    W, H = desc["width"], desc["height"]
    A = O.step(desc, A, use_fft=use_fft)  # warm-up step
    t0 = time.perf_counter()
    n = 0
    while True:
        A = O.step(desc, A, use_fft=use_fft)
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": round(W * H * n / dt / 1e6, 3), "unit": "Mcell-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"FP64 oracle {'FFT' if use_fft else 'direct'} step (oracle/lenia_oracle.c), "
                      f"{n} steps of the same {W}x{H} K={len(desc['kernels'])} world after 1 "
                      f"warm-up step, OpenMP {threads} threads, {dt:.1f} s"}


def cpu_baseline(desc: dict, budget_s: float = 12.0) -> dict:
    import oracle as O

    W, H = desc["width"], desc["height"]
    x0, y0, pw, ph = desc["initRect"]
    A = O.init_patch(W, H, O.module_seed(7777), x0, y0, pw, ph)
    out = cpu_sample(desc, A, True, budget_s)
    threads = O.max_threads()
    O.set_threads(1)
    try:
        one = cpu_sample(desc, A, True, budget_s / 3)
    finally:
        O.set_threads(threads)
    out["one_core"] = {"value": one["value"], "cores": 1, "sample": one["sample"]}
    return out


def live_field_bench(ctx, steps: int) -> dict:
    """The headline kernels (three R=24 rings, same growth) in growth mode 1
    from the oracle's value-noise field: every cell stays inside (0, 1), so
    this is the step rate on a live field, beside the headline world that is
    empty after 18 steps."""
    import numpy as np

    import lenia
    import oracle as O
    from lenia_cases import config1_live, live_cells

    desc = config1_live()
    W, H = desc["width"], desc["height"]
    A0 = O.init_value_noise(W, H, O.module_seed(7777), 0.42, 26.0).astype(np.float32)
    with lenia.World(ctx, desc) as w:
        w.write(A0)
        w.autotune(20)
        w.write(A0)
        w.step(50)
        ctx.sync()
        w.step(steps)
        ms = ctx.last_gpu_ms()
        A = w.read()
    return {"workload": "1024x512 three R=24 rings, growth mode 1, value-noise field",
            "steps": steps, "ms_per_step": round(ms / steps, 5),
            "mcell_steps_per_s": round(W * H * steps / (ms * 1e-3) / 1e6, 1),
            "live_cell_frac": round(live_cells(A) / A.size, 4), "state_std": round(float(A.std()), 4)}


def tune_sweep_plans(ctx) -> dict:
    """radix plans for the sweep's side lengths, each tuned on a square world of
    that side with two R=8 ring kernels (lenia_plan_autotune); the sweep's
    worlds are created afterwards and take them from the plan cache"""
    import lenia
    from lenia_cases import FFT, INIT_PATCH, SWEEP_SIDES, ring

    out = {}
    for n in SWEEP_SIDES:
        d = dict(width=n, height=n, convMode=FFT, growthMode=0, dt=0.1, radius=8,
                 kernels=[ring(8, [1.0], 0.2, 0.03, 0.5), ring(8, [1.0, 0.5], 0.3, 0.05, 0.5)],
                 initMode=INIT_PATCH, initRect=(0, 0, 0, 0))
        pw, ph, _ = lenia.plan_autotune(ctx, d, 6)
        out[str(n)] = {"rows": hex(pw), "cols": hex(ph)}
    return out


def sweep_bench(ctx, steps: int = 1000, graph_steps: int = 50, plain_steps: int = 64,
                tune_plans: bool = True, waves=None) -> dict:
    """256 uneven worlds stepped in one batch (graph-replayed) vs the same worlds
    stepped one world per launch sequence, one after another (the PresetLab
    loop, Apps/LifePresetLab/main.cpp:718-806).  The plain loop replays each
    world's own 32-step graphs (plain_steps is a multiple of 32).  Both sides
    are timed the same way, by the host clock between two synchronisations;
    the batch's GPU-event time is reported too.  waves: fixed (row, column)
    executors (lenia_batch_set_waves) instead of the autotuner's choice."""
    import lenia
    from lenia_cases import sweep_worlds

    graph_steps = min(graph_steps, steps)
    if steps % graph_steps:
        raise SystemExit(f"bench.py: sweep steps ({steps}) must be a multiple of {graph_steps}")
    if plain_steps % 32:
        raise SystemExit("bench.py: --plain-steps must be a multiple of 32 (graph replay)")
    plans = tune_sweep_plans(ctx) if tune_plans else None
    descs = sweep_worlds(256)
    worlds = [lenia.World(ctx, d) for d in descs]
    try:
        for i, w in enumerate(worlds):
            w.reset(1000 + i)
            w.frame(0, 1)
        cells = sum(d["width"] * d["height"] for d in descs)
        bytes_step = sum(d["width"] * d["height"] *
                         fft_bytes_per_cell(d["width"], len(d["kernels"]))["step"] for d in descs)
        with lenia.Batch(ctx, worlds) as b:
            tuned_ms = b.autotune(10)  # launch shapes for this batch (bits unchanged)
            if waves is not None:
                b.set_waves(*waves)
            b.step(graph_steps, graph_steps)  # warm-up + graph capture
            ctx.sync()
            t0 = time.perf_counter()
            b.step(steps, graph_steps)
            ctx.sync()
            batch_wall = time.perf_counter() - t0
            batch_ms = ctx.last_gpu_ms()
            # the same batch with every world on the workgroup-LDS executor,
            # timed the same way (the executors give the same bits)
            tuned_waves = b.waves()
            b.set_waves(0, 0)
            b.step(graph_steps, graph_steps)
            b.step(steps, graph_steps)
            ctx.sync()
            lds_ms = ctx.last_gpu_ms()
            b.set_waves(*tuned_waves)
        batch_rate = cells * steps / (batch_ms * 1e-3) / 1e6
        batch_wall_rate = cells * steps / batch_wall / 1e6
        plain = None
        if plain_steps:
            for w in worlds:
                w.step(32)  # each world captures its own 32-step graph
            ctx.sync()
            t0 = time.perf_counter()
            for w in worlds:
                w.step(plain_steps)
            ctx.sync()
            plain_s = time.perf_counter() - t0
            plain_rate = cells * plain_steps / plain_s / 1e6
            plain = {"steps_per_world": plain_steps, "graph_replayed": True,
                     "wall_mcell_steps_per_s": round(plain_rate, 1),
                     "speedup_batched_vs_plain_wall": round(batch_wall_rate / plain_rate, 2)}
        gbs = bytes_step * steps / (batch_ms * 1e-3) / 1e9
        return {"worlds": len(descs), "cells": cells, "steps": steps,
                "graph_steps": graph_steps, "ms_per_step": round(batch_ms / steps, 5),
                "mcell_steps_per_s": round(batch_rate, 1),
                "wall_mcell_steps_per_s": round(batch_wall_rate, 1),
                "plain": plain,
                "speedup_vs_plain": plain["speedup_batched_vs_plain_wall"] if plain else None,
                "achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                "bytes_per_step": int(bytes_step), "autotuned_ms_per_step": round(tuned_ms, 5),
                "executors": {"tuned_waves_rows_cols": list(tuned_waves),
                              "lds_only_ms_per_step": round(lds_ms / steps, 5)},
                "radix_plans": plans}
    finally:
        for w in worlds:
            w.close()


def adjoint_bench(ctx, T: int = 64, ckpt: int = 8, reps: int = 20) -> dict:
    """differentiable-rollout world (256x256, two ring kernels R=16): wall time
    of the T-step forward rollout (which keeps checkpoints every `ckpt` steps)
    and of its reverse pass (segment recompute included), host arrays in and
    out, best of reps after a second of warm-up calls (the clocks of a GPU that
    was idle ramp up over the first calls); also the reverse pass that reruns
    the forward sweep"""
    import numpy as np

    import lenia
    from lenia_cases import INIT_NONE, config4

    desc = config4()
    with lenia.World(ctx, desc) as w0:
        w0.reset(77)
        w0.frame(0)
        A0 = w0.read()
    rng = np.random.default_rng(5)
    gT = rng.standard_normal(A0.shape).astype(np.float32)
    with lenia.World(ctx, dict(desc, initMode=INIT_NONE)) as w:
        w.write(A0)
        w.autotune(10)
        w.rollout(A0, T, ckpt)
        w.rollout_grad(None, T, gT, ckpt)  # workspace, graphs, spectra
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < 1.0:
            w.rollout(A0, T, ckpt)
            w.rollout_grad(None, T, gT, ckpt)
        fwd, bwd, full = [], [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            w.rollout(A0, T, ckpt)  # keeps the checkpoints
            t1 = time.perf_counter()
            w.rollout_grad(None, T, gT, ckpt)  # recompute + reverse steps
            t2 = time.perf_counter()
            w.rollout_grad(A0, T, gT, ckpt)  # forward sweep + recompute + reverse
            t3 = time.perf_counter()
            fwd.append(t1 - t0)
            bwd.append(t2 - t1)
            full.append(t3 - t2)
    f, b, fb = min(fwd) * 1e3, min(bwd) * 1e3, min(full) * 1e3
    return {"world": "256x256 K=2 R=16 FFT", "T": T, "ckpt_every": ckpt,
            "forward_ms": round(f, 4), "backward_ms": round(b, 4),
            "forward_plus_adjoint_over_forward": round((f + b) / f, 3),
            "median_forward_ms": round(float(np.median(fwd)) * 1e3, 4),
            "median_backward_ms": round(float(np.median(bwd)) * 1e3, 4),
            "backward_with_forward_sweep_ms": round(fb, 4),
            "mcell_steps_per_s_forward": round(256 * 256 * T / f / 1e3, 1)}


def direct_forms_bench(ctx, steps: int = 200) -> dict:
    """1280x720, one exp-bump kernel R=32, full-frame seeded noise: per-step
    kernel time of the FP32 vector direct form, the BF16 MFMA forms (plain and
    split) and the FFT form.  Algorithmic flops = 2 x nnz per cell-step, nnz =
    #{(dx, dy): 0 < dx^2 + dy^2 < R^2} (the exp-bump core vanishes at r = 0 and
    r >= 1); executed MFMA flops include the band's zeros and the split's
    extra products."""
    import numpy as np

    import lenia
    from lenia_cases import BF16, BF16X3, DIRECT, FFT, config2

    R = 32
    d = np.arange(-R, R + 1)
    r2 = d[:, None] ** 2 + d[None, :] ** 2
    nnz = int(np.count_nonzero((r2 > 0) & (r2 < R * R)))
    ksteps = (32 + 2 * R + 15) // 16
    out = {"world": "1280x720 K=1 exp-bump R=32", "nnz": nnz, "forms": {}}
    for conv, name, mfmas in ((DIRECT, "fp32_direct", 0), (BF16, "bf16_mfma", 1),
                              (BF16X3, "bf16x3_mfma", 4), (FFT, "fp32_fft", 0)):
        desc = config2(conv=conv)
        cells = desc["width"] * desc["height"]
        with lenia.World(ctx, desc) as w:
            w.reset(7777)
            w.frame(0, 1)
            w.step(10)
            ms = sum(w.pass_times(steps))
        rec = {"us_per_step": round(ms * 1e3, 3),
               "mcell_steps_per_s": round(cells / (ms * 1e-3) / 1e6, 1)}
        alg_tflops = 2.0 * nnz * cells / (ms * 1e-3) / 1e12
        if conv == DIRECT:
            rec["algorithmic_tflops"] = round(alg_tflops, 2)
            rec["valu_fp32_frac"] = round(alg_tflops / VALU_FP32_PEAK_TFLOPS, 4)
        elif mfmas:
            exe = 2.0 * 16 * ksteps * (2 * R + 1) * mfmas * cells / (ms * 1e-3) / 1e12
            rec["algorithmic_tflops"] = round(alg_tflops, 2)
            rec["executed_mfma_tflops"] = round(exe, 2)
            rec["mfma_util_algorithmic"] = round(alg_tflops / MFMA_BF16_PEAK_TFLOPS, 4)
            rec["mfma_util_executed"] = round(exe / MFMA_BF16_PEAK_TFLOPS, 4)
        out["forms"][name] = rec
    return out


def matrix_bench(ctx, cpu_s: float, tune_plans: bool = True) -> dict:
    """The configurations other than the headline, each at its stated length:
    512^2 direct FP32 R=13 (200 steps), and the headline kernels' FFT form at
    1920x1080 and 4096^2 (500 steps, launch shapes autotuned).  Rates from HIP
    events around the timed steps; per-pass times from PASS_SAMPLES-step
    per-kernel event samples; a bounded FP64 oracle sample beside each
    (cpu_s seconds; 0 skips them)."""
    import lenia
    import oracle as O
    from lenia_cases import config0, config1

    out = {}
    # 512^2, direct FP32, one exp-bump kernel R=13
    d = config0()
    W, H = d["width"], d["height"]
    nnz = window_nnz(d["kernels"][0])
    with lenia.World(ctx, d) as w:
        w.reset(7777)
        w.frame(0, 1)
        w.step(64)  # also records the world's step graph, outside the timed steps
        ctx.sync()
        w.step(200)
        ms = ctx.last_gpu_ms()
        step_ms = w.pass_times(PASS_SAMPLES)[0]
    tfl = 2.0 * nnz * W * H / (step_ms * 1e-3) / 1e12
    rec = {"workload": "lenia-direct-512x512-R13 (FP32 vector direct)", "steps": 200,
           "ms_per_step": round(ms / 200, 5),
           "mcell_steps_per_s": round(W * H * 200 / (ms * 1e-3) / 1e6, 1),
           "roofline": {"bound": "valu", "kernel": "k_direct", "ms": round(step_ms, 5),
                        "nnz": nnz, "algorithmic_flops_per_cell": 2 * nnz,
                        "achieved": round(tfl, 2), "peak": VALU_FP32_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(tfl / VALU_FP32_PEAK_TFLOPS, 4)}}
    if cpu_s:
        x0, y0, pw, ph = d["initRect"]
        A = O.init_patch(W, H, O.module_seed(7777), x0, y0, pw, ph)
        rec["cpu"] = cpu_sample(d, A, False, cpu_s)
    out["config0_512x512_direct_R13"] = rec
    # the headline kernels' FFT form at the larger sizes
    for W, H in ((1920, 1080), (4096, 4096)):
        d = config1(W, H)
        K = len(d["kernels"])
        if tune_plans:
            lenia.plan_autotune(ctx, d, 6)  # radix plans of this width and height
        with lenia.World(ctx, d) as w:
            w.reset(7777)
            w.frame(0, 1)
            w.autotune(10)
            w.step(64)  # also records the world's step graph, outside the timed steps
            ctx.sync()
            w.step(500)
            ms = ctx.last_gpu_ms()
            passes = w.pass_times(PASS_SAMPLES)
            tiles = w.resolved()
        roof = fft_roofline(W, H, K, passes, (tiles["colEngine"], tiles["rowEngine"]))
        dom = max(roof, key=lambda k: roof[k]["ms"])
        step_gbs = fft_bytes_per_cell(W, K)["step"] * W * H / (ms / 500 * 1e-3) / 1e9
        rec = {"workload": f"lenia-fft-{W}x{H}-K3-R24", "steps": 500,
               "ms_per_step": round(ms / 500, 5),
               "mcell_steps_per_s": round(W * H * 500 / (ms * 1e-3) / 1e6, 1),
               "tiles": tiles, "passes": roof,
               "roofline": dict(roof[dom], bound="hbm", peak=HBM_PEAK_GBS, unit="GB/s",
                                achieved=roof[dom]["achieved_gbs"]),
               "step_achieved_gbs": round(step_gbs, 1),
               "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4)}
        if cpu_s and W * H <= 1920 * 1080:
            x0, y0, pw, ph = d["initRect"]
            A = O.init_patch(W, H, O.module_seed(7777), x0, y0, pw, ph)
            rec["cpu"] = cpu_sample(d, A, True, cpu_s)
        out[f"config1_{W}x{H}_fft"] = rec
    return out


def _scene_text(name: str, resolution=None, edit=None) -> str:
    with open(os.path.join(SCENES, name)) as f:
        j = json.load(f)
    if resolution:
        j["resolution"] = list(resolution)
    if edit:
        edit(j["modules"][0]["params"])
    return json.dumps(j)


def lifebench_bench(ctx, warmup: int = 30, frames: int = 240) -> dict:
    """The LifeBench runs of the reference's FFT acceptance list
    (docs/specs/phase9_lenia_fft.md:165-168) through the scene runner: (a)
    Presets/lenia_basic.json at 1280x720 (direct, R=13), (b) the same scene
    with radius 32, (c) Presets/lenia_multikernel.json (FFT, R=24, K=3, sim
    1024x512).  LifeBench's defaults: 30 warm-up frames, 240 timed; GPU time
    per frame covers the world steps and the colour layer (the reference's
    frame time also holds its compositing and output passes).  (c) also runs
    300 frames twice from its seed and compares the final layers' md5."""
    import lenia

    ref_ms = {"b_direct_R32_1280x720": 19.4, "c_multikernel_fft_1024x512": 1.13}
    cases = [("a_basic_direct_R13_1280x720", _scene_text("lenia_basic.json", (1280, 720))),
             ("b_direct_R32_1280x720",
              _scene_text("lenia_basic.json", (1280, 720), lambda p: p.update(radius=32))),
             ("c_multikernel_fft_1024x512", _scene_text("lenia_multikernel.json"))]
    out = {}
    for name, text in cases:
        with lenia.Scene(ctx, text) as sc:
            sc.step(warmup)
            ctx.sync()
            sc.step(frames)
            ms = ctx.last_gpu_ms() / frames
            w = sc.world(0)
            rec = {"gpu_ms_per_frame": round(ms, 5), "frames": frames,
                   "sim": f"{w.W}x{w.H}", "kernels": len(w.desc["kernels"]),
                   "convMode": ["direct", "fft", "bf16", "bf16x3"][int(w.desc["convMode"])]}
        if name in ref_ms:
            rec["reference_m5max_ms_per_frame"] = ref_ms[name]
            rec["reference_over_this"] = round(ref_ms[name] / ms, 1)
        out[name] = rec
    md5 = []
    for _ in range(2):
        with lenia.Scene(ctx, cases[2][1]) as sc:
            sc.step(300)
            md5.append(hashlib.md5(sc.layer(0).tobytes()).hexdigest())
    out["c_multikernel_300_frames_md5"] = md5[0]
    out["c_multikernel_300_frames_deterministic"] = md5[0] == md5[1]
    return out


def copy_calibration(mib: int, reps: int = 10) -> dict:
    """a plain streaming kernel of known size (torch.neg over `mib` MiB of fp32:
    reads and writes mib MiB each) whose FETCH_SIZE / WRITE_SIZE counters scale
    the counter readings of the Lenia kernels to bytes"""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.rand(n, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    return {"calibration": "torch.neg fp32", "bytes_read": n * 4, "bytes_written": n * 4,
            "launches": reps}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_fft_rows_lenia_fft_1024x512_K3_traffic.md")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy calibration) from the committed counter table, or None"""
    try:
        with open(TRAFFIC_FILE) as f:
            lines = f.read().splitlines()
    except OSError:
        return None
    for line in lines:
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 9 and cells[0].startswith(kernel + " "):
            try:
                return int(cells[6])
            except ValueError:
                return None
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline samples")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-sweep", action="store_true", help="skip the 256-world batched sweep")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the rollout-gradient leg")
    ap.add_argument("--no-direct", action="store_true", help="skip the 1280x720 direct-forms leg")
    ap.add_argument("--no-matrix", action="store_true", help="skip the other configurations")
    ap.add_argument("--no-lifebench", action="store_true", help="skip the LifeBench scenes")
    ap.add_argument("--no-live", action="store_true", help="skip the live-field rate")
    ap.add_argument("--no-plan-tune", action="store_true",
                    help="keep the built-in radix plans (no lenia_plan_autotune)")
    ap.add_argument("--engines", default="", metavar="ROW,COL",
                    help="pass executors of the headline world: lds, wave or split (with --tiles: "
                         "a wave pass takes its threads from ROWT / COLT and needs ROWP / COLC 0)")
    ap.add_argument("--plans", default="", metavar="PLANW,PLANH",
                    help="radix plans of the headline world (hex, as in the descriptor); implies "
                         "no plan tuning")
    ap.add_argument("--tiles", default="", metavar="ROWT,ROWP,COLT,COLC",
                    help="fixed launch shapes of the headline world instead of autotuning "
                         "(row threads, row pairs per tile, column threads, columns per tile)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain fp32 streaming kernel over MIB MiB (counter scaling)")
    ap.add_argument("--only", choices=["sweep", "adjoint", "direct", "matrix", "lifebench", "live"],
                    default="", help="run one side leg alone and print its JSON object (profiling)")
    ap.add_argument("--sweep-steps", type=int, default=1000, help="timed batched sweep steps")
    ap.add_argument("--sweep-waves", default="", metavar="ROW,COL",
                    help="fixed executors of the sweep batch (waves per workgroup, 0 = LDS, "
                         "3 = kernel-split) instead of the autotuner's choice")
    ap.add_argument("--plain-steps", type=int, default=64,
                    help="steps per world of the one-world-per-launch loop, a multiple of 32 "
                         "(0: skip it)")
    args = ap.parse_args()
    if args.copy_calibration:
        print(json.dumps(copy_calibration(args.copy_calibration)))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(worlds shard across GPUs with no communication; see the sharding notes)",
              file=sys.stderr)
        return 2
    cpu_s = 0.0 if args.no_cpu else max(1.0, args.cpu_seconds / 4)
    if args.only:
        import lenia
        sw = tuple(int(x) for x in args.sweep_waves.split(",")) if args.sweep_waves else None
        legs = {"sweep": lambda c: sweep_bench(c, args.sweep_steps, plain_steps=args.plain_steps, waves=sw,
                                               tune_plans=not args.no_plan_tune),
                "adjoint": adjoint_bench, "direct": direct_forms_bench,
                "matrix": lambda c: matrix_bench(c, cpu_s, not args.no_plan_tune),
                "lifebench": lifebench_bench,
                "live": lambda c: live_field_bench(c, args.steps)}
        with lenia.Context(0) as ctx:
            print(json.dumps({args.only: legs[args.only](ctx)}))
        return 0

    import lenia
    from lenia_cases import config1

    desc = config1()
    if args.tiles:
        rt, rp, ct, cc = (int(x) for x in args.tiles.split(","))
        desc.update(rowThreads=rt, rowPairsPerTile=rp, colThreads=ct, colsPerTile=cc)
    if args.engines:
        codes = dict({v: k for k, v in ENGINE_NAMES.items()}, split=ENGINE_WAVE_SPLIT)
        re_, ce = args.engines.split(",")
        desc.update(rowEngine=codes[re_], colEngine=codes[ce])
    if args.plans:
        pw, ph = (int(x, 16) for x in args.plans.split(","))
        desc.update(fftPlanW=pw, fftPlanH=ph)
        args.no_plan_tune = True
    W, H, K = desc["width"], desc["height"], len(desc["kernels"])
    cells = W * H
    with lenia.Context(0) as ctx:
        if not args.tiles and not args.no_plan_tune and not args.plans:
            # radix plans of the headline world's width and height (plan cache)
            lenia.plan_autotune(ctx, desc, 10)
        w = lenia.World(ctx, desc)
        w.reset(7777)
        w.frame(0, 1)
        if not (args.tiles or args.engines):
            w.autotune(20)  # launch shapes for this world (bits unchanged)
        w.step(args.warmup)
        ctx.sync()
        t0 = time.perf_counter()
        w.step(args.steps)
        ctx.sync()
        wall = time.perf_counter() - t0
        gpu_ms = ctx.last_gpu_ms()
        # per-pass durations: a fixed sample, whatever --steps is
        passes = w.pass_times(PASS_SAMPLES)
        tiles = w.resolved()
        A = w.read()
        w.close()
        live = None if args.no_live else live_field_bench(ctx, args.steps)
        sw = tuple(int(x) for x in args.sweep_waves.split(",")) if args.sweep_waves else None
        batched = None if args.no_sweep else sweep_bench(ctx, args.sweep_steps,
                                                          plain_steps=args.plain_steps,
                                                          tune_plans=not args.no_plan_tune, waves=sw)
        adjoint = None if args.no_adjoint else adjoint_bench(ctx)
        direct = None if args.no_direct else direct_forms_bench(ctx)
        matrix = None if args.no_matrix else matrix_bench(ctx, cpu_s, not args.no_plan_tune)
        lifebench = None if args.no_lifebench else lifebench_bench(ctx)
    sec = gpu_ms / 1e3
    value = cells * args.steps / sec / 1e6
    roof = fft_roofline(W, H, K, passes, (tiles["colEngine"], tiles["rowEngine"]))
    dom = max(roof, key=lambda k: roof[k]["ms"])
    step_gbs = fft_bytes_per_cell(W, K)["step"] * cells * args.steps / sec / 1e9
    out = {
        "metric": "Mcell-steps/s, Lenia 3-kernel R=24 FFT 1024x512; batched & adjoint vs plain",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(gpu_ms / args.steps, 5),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 256x256 uniform patch (rand01 PCG hash), scene seed 7777",
        "config": {"workload": "lenia-fft-1024x512-K3-R24", "width": W, "height": H,
                   "kernels": K, "radius": 24, "growthMode": 0, "dt": 0.1,
                   "fftPlanW": tiles["fftPlanW"], "fftPlanH": tiles["fftPlanH"],
                   "rowPairsPerTile": tiles["rowPairsPerTile"],
                   "colsPerTile": tiles["colsPerTile"], "rowThreads": tiles["rowThreads"],
                   "colThreads": tiles["colThreads"],
                   "rowEngine": ENGINE_NAMES[tiles["rowEngine"]],
                   "colEngine": ENGINE_NAMES[tiles["colEngine"]],
                   "autotuned": not (args.tiles or args.engines),
                   "radix_plans_tuned": not (args.tiles or args.no_plan_tune)},
        "passes": roof,
        "pass_samples": PASS_SAMPLES,
        "roofline": {"bound": "hbm", "kernel": roof[dom]["kernel"],
                     "achieved": roof[dom]["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": roof[dom]["frac"],
                     "traffic": measured_traffic(roof[dom]["kernel"]),
                     "ms": roof[dom]["ms"],
                     "algorithmic_bytes_per_launch": roof[dom]["algorithmic_bytes_per_launch"],
                     "algorithmic_bytes_per_cell": roof[dom]["algorithmic_bytes_per_cell"],
                     "step_achieved_gbs": round(step_gbs, 1),
                     "step_bytes_per_cell": round(fft_bytes_per_cell(W, K)["step"], 3)},
        "state_mean": float(A.mean()),
    }
    for key, val in (("live_field", live), ("batched", batched), ("adjoint", adjoint), ("direct_forms", direct),
                     ("matrix", matrix), ("lifebench", lifebench)):
        if val is not None:
            out[key] = val
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(desc, args.cpu_seconds)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
