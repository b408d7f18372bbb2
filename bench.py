"""Lenia field-update benchmark on one MI355X (the LifeBench measurement,
Apps/LifeBench/main.cpp:43-177, for the Lenia hot path).

Workload (the headline metric's configuration): FFT-form Lenia, 1024x512 FP32
toroidal field, three ring kernels R=24 with shell peaks [1], [1,.5],
[.5,1,.25], weights .5/.3/.2, growth (mu, sigma) (.15,.015) (.26,.036)
(.12,.012), dt .1, seeded 256x256 patch init.  A step is one full field
update; the field stays resident in HBM.

Prints one JSON line:  value = Mcell-steps/s = W*H*steps / GPU seconds.
  roofline     dominant kernel: algorithmic bytes per launch / its mean
               launch time (HIP events on the library's stream)
  cpu_baseline the FP64 oracle's FFT step on the host cores, bounded sample
"""
from __future__ import annotations

import argparse
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


def cpu_baseline(desc: dict, budget_s: float = 12.0) -> dict:
    import numpy as np

    import oracle as O

    threads = O.max_threads()
    W, H = desc["width"], desc["height"]
    m = O.module_seed(7777)
    x0, y0, pw, ph = desc["initRect"]
    A = O.init_patch(W, H, m, x0, y0, pw, ph)
    A = O.step(desc, A, use_fft=True)  # warm-up step
    t0 = time.perf_counter()
    n = 0
    while True:
        A = O.step(desc, A, use_fft=True)
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    # the same step on one core (a shorter sample)
    O.set_threads(1)
    t1 = time.perf_counter()
    n1 = 0
    while True:
        A = O.step(desc, A, use_fft=True)
        n1 += 1
        if time.perf_counter() - t1 >= budget_s / 3:
            break
    dt1 = time.perf_counter() - t1
    O.set_threads(threads)
    return {"value": round(W * H * n / dt / 1e6, 3), "unit": "Mcell-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"FP64 oracle FFT step (oracle/lenia_oracle.c), {n} steps of the same "
                      f"{W}x{H} K={len(desc['kernels'])} world after 1 warm-up step, "
                      f"OpenMP {threads} threads, {dt:.1f} s",
            "one_core": {"value": round(W * H * n1 / dt1 / 1e6, 3), "cores": 1,
                         "sample": f"{n1} steps on 1 thread, {dt1:.1f} s"}}


def sweep_bench(ctx, steps: int = 1000, graph_steps: int = 50, plain_steps: int = 20) -> dict:
    """256 uneven worlds stepped in one batch (graph-replayed) vs the same worlds
    stepped one world per launch sequence, one after another (the PresetLab
    loop, Apps/LifePresetLab/main.cpp:718-806)."""
    import lenia
    from lenia_cases import sweep_worlds

    graph_steps = min(graph_steps, steps)
    if steps % graph_steps:
        raise SystemExit(f"bench.py: sweep steps ({steps}) must be a multiple of {graph_steps}")
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
            b.step(graph_steps, graph_steps)  # warm-up + graph capture
            ctx.sync()
            b.step(steps, graph_steps)
            batch_ms = ctx.last_gpu_ms()
        batch_rate = cells * steps / (batch_ms * 1e-3) / 1e6
        plain_rate = None
        if plain_steps:
            # plain: each world alone, one after the other
            for w in worlds:
                w.step(2)
            ctx.sync()
            t0 = time.perf_counter()
            for w in worlds:
                w.step(plain_steps)
            ctx.sync()
            plain_s = time.perf_counter() - t0
            plain_rate = cells * plain_steps / plain_s / 1e6
        gbs = bytes_step * steps / (batch_ms * 1e-3) / 1e9
        return {"worlds": len(descs), "cells": cells, "steps": steps,
                "graph_steps": graph_steps, "ms_per_step": round(batch_ms / steps, 5),
                "mcell_steps_per_s": round(batch_rate, 1),
                "plain_mcell_steps_per_s": round(plain_rate, 1) if plain_rate else None,
                "speedup_vs_plain": round(batch_rate / plain_rate, 2) if plain_rate else None,
                "achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                "bytes_per_step": int(bytes_step), "autotuned_ms_per_step": round(tuned_ms, 5)}
    finally:
        for w in worlds:
            w.close()


def adjoint_bench(ctx, T: int = 64, ckpt: int = 8, reps: int = 5) -> dict:
    """differentiable-rollout world (256x256, two ring kernels R=16): wall time
    of the T-step forward rollout (which keeps checkpoints every `ckpt` steps)
    and of its reverse pass (segment recompute included), host arrays in and
    out, best of reps; also the reverse pass that reruns the forward sweep"""
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
            "backward_with_forward_sweep_ms": round(fb, 4),
            "mcell_steps_per_s_forward": round(256 * 256 * T / f / 1e3, 1)}


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak, published
VALU_FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector peak, published


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
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy calibration), or None"""
    try:
        with open(TRAFFIC_FILE) as f:
            text = f.read()
    except OSError:
        return None
    vals = {}
    for line in text.splitlines():
        if ":" in line and line.strip().startswith("- "):
            k, v = line.strip()[2:].split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        return int(vals[kernel + " hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-sweep", action="store_true", help="skip the 256-world batched sweep")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the rollout-gradient leg")
    ap.add_argument("--no-direct", action="store_true", help="skip the 1280x720 direct-forms leg")
    ap.add_argument("--tiles", default="", metavar="ROWT,ROWP,COLT,COLC",
                    help="fixed launch shapes of the headline world instead of autotuning "
                         "(row threads, row pairs per tile, column threads, columns per tile)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain fp32 streaming kernel over MIB MiB (counter scaling)")
    ap.add_argument("--only", choices=["sweep", "adjoint", "direct"], default="",
                    help="run one side leg alone and print its JSON object (for profiling)")
    ap.add_argument("--sweep-steps", type=int, default=1000, help="timed batched sweep steps")
    ap.add_argument("--plain-steps", type=int, default=20,
                    help="steps per world of the one-world-per-launch loop (0: skip it)")
    args = ap.parse_args()
    if args.copy_calibration:
        print(json.dumps(copy_calibration(args.copy_calibration)))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(worlds shard across GPUs with no communication; see the sharding notes)",
              file=sys.stderr)
        return 2
    if args.only:
        import lenia
        legs = {"sweep": lambda c: sweep_bench(c, args.sweep_steps, plain_steps=args.plain_steps),
                "adjoint": adjoint_bench, "direct": direct_forms_bench}
        with lenia.Context(0) as ctx:
            print(json.dumps({args.only: legs[args.only](ctx)}))
        return 0

    import lenia
    from lenia_cases import config1

    desc = config1()
    if args.tiles:
        rt, rp, ct, cc = (int(x) for x in args.tiles.split(","))
        desc.update(rowThreads=rt, rowPairsPerTile=rp, colThreads=ct, colsPerTile=cc)
    W, H, K = desc["width"], desc["height"], len(desc["kernels"])
    cells = W * H
    with lenia.Context(0) as ctx, lenia.World(ctx, desc) as w:
        w.reset(7777)
        w.frame(0, 1)
        if not args.tiles:
            w.autotune(20)  # launch shapes for this world (bits unchanged)
        w.step(args.warmup)
        ctx.sync()
        t0 = time.perf_counter()
        w.step(args.steps)
        ctx.sync()
        wall = time.perf_counter() - t0
        gpu_ms = ctx.last_gpu_ms()
        # per-pass durations, events between kernels on the library stream
        passes = w.pass_times(min(args.steps, 500))
        tiles = w.resolved()
        A = w.read()
        batched = None if args.no_sweep else sweep_bench(ctx, args.sweep_steps,
                                                          plain_steps=args.plain_steps)
        adjoint = None if args.no_adjoint else adjoint_bench(ctx)
        direct = None if args.no_direct else direct_forms_bench(ctx)
    sec = gpu_ms / 1e3
    value = cells * args.steps / sec / 1e6
    bpc = fft_bytes_per_cell(W, K)
    names = ["cols", "rows"]
    per = {names[i]: passes[i] for i in range(len(passes))}
    dom = max(per, key=per.get)
    dom_kernel = "k_fft_rows" if dom == "rows" else "k_fft_cols"
    achieved = bpc[dom] * cells / (per[dom] * 1e-3) / 1e9
    step_gbs = bpc["step"] * cells * args.steps / sec / 1e9
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
                   "colThreads": tiles["colThreads"], "autotuned": not args.tiles},
        "passes_ms": {k: round(v, 5) for k, v in per.items()},
        "roofline": {"bound": "hbm", "kernel": dom_kernel,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom_kernel),
                     "algorithmic_bytes_per_launch": int(round(bpc[dom] * cells)),
                     "algorithmic_bytes_per_cell": round(bpc[dom], 3),
                     "step_achieved_gbs": round(step_gbs, 1),
                     "step_bytes_per_cell": round(bpc["step"], 3)},
        "state_mean": float(A.mean()),
    }
    if batched is not None:
        out["batched"] = batched
    if adjoint is not None:
        out["adjoint"] = adjoint
    if direct is not None:
        out["direct_forms"] = direct
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(desc, args.cpu_seconds)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
