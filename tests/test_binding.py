"""LeniaHipModule (life-simulation_amd/binding): the module class a maintainer
registers in place of LeniaModule (Modules/RegisterModules.cpp:23), compiled
with g++ against include/lenia_mi355x.h together with a driver that steps it
the way SceneRunner::step steps a module (SceneRunner.cpp:146-215).

CPU: it builds, and setup() resolves every reference preset's module as the
library's scene layer does (LeniaModule::setup, Lenia.cpp:208-408).
GPU: stepped through encode() with per-frame ParameterBus values, its field
and colour layer equal the scene runner's byte for byte."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENES = os.path.join(ROOT, "tests", "golden", "scenes")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    from lenia import _abi

    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no C++ compiler")
    exe = str(tmp_path_factory.mktemp("binding") / "binding_driver")
    lib = _abi.LIB_PATH
    subprocess.run([cxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                    "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "life-simulation_amd", "binding"),
                    "-I", os.path.join(ROCM, "include"),
                    os.path.join(ROOT, "tests", "binding_driver.cpp"),
                    os.path.join(ROOT, "life-simulation_amd", "binding", "lenia_hip_module.cpp"),
                    "-o", exe, lib, "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
                    "-Wl,-rpath," + os.path.dirname(lib), "-Wl,-rpath," + os.path.join(ROCM, "lib")],
                   check=True)
    return exe


def _module(path):
    scene = json.load(open(path))
    W, H = scene["resolution"]
    mods = [m for m in scene["modules"] if m.get("type") == "Lenia" and m.get("enabled", True)]
    return scene, mods, W, H


def _scene_files():
    out = []
    for f in sorted(os.listdir(SCENES)):
        if not f.endswith(".json"):
            continue
        scene = json.load(open(os.path.join(SCENES, f)))
        if isinstance(scene, dict) and "modules" in scene and any(
                m.get("type") == "Lenia" for m in scene["modules"]):
            out.append(f)
    return out


@pytest.mark.parametrize("fname", _scene_files())
def test_setup_resolves_like_the_scene_layer(driver, tmp_path, fname):
    import lenia

    path = os.path.join(SCENES, fname)
    scene, mods, W, H = _module(path)
    text = open(path).read()
    cwd = os.getcwd()
    os.chdir(SCENES)  # organism paths are relative to the working directory
    try:
        for i, m in enumerate(mods):
            want, cmap, name = lenia.describe_scene(text, i)
            pj = tmp_path / f"params{i}.json"
            pj.write_text(json.dumps(m["params"]))
            out = subprocess.run([driver, "describe", str(pj), m["name"], str(W), str(H)], check=True,
                                 capture_output=True, text=True).stdout
            got = {}
            for line in out.splitlines():
                f = line.split()
                got[f[0] if f[0] != "kernel" else f"kernel{f[1]}"] = f[1:] if f[0] == "kernel" else f[1:]
            assert int(got["width"][0]) == want["width"] and int(got["height"][0]) == want["height"]
            assert int(got["convMode"][0]) == want["convMode"]
            assert int(got["nKernels"][0]) == want["nKernels"]
            assert int(got["radius"][0]) == want["radius"]
            assert np.float32(got["dt"][0]) == np.float32(want["dt"])
            assert int(got["growthMode"][0]) == want["growthMode"]
            assert int(got["initMode"][0]) == want["initMode"]
            for k in range(want["nKernels"]):
                kk = want["kernels"][k]
                g = got[f"kernel{k}"]
                assert np.float32(g[1]) == np.float32(kk["radius"])
                assert np.float32(g[2]) == np.float32(kk["mu"]) and np.float32(g[3]) == np.float32(kk["sigma"])
                assert np.float32(g[4]) == np.float32(kk["weight"]) and int(g[5]) == len(kk["betas"])
            assert [np.float32(v) for v in got["palette_d"]] == [np.float32(v) for v in cmap["d"]]
    finally:
        os.chdir(cwd)


@pytest.fixture(scope="module")
def ctx():
    import lenia

    with lenia.Context(0) as c:
        yield c


@pytest.mark.gpu
@pytest.mark.parametrize("fname,bus,frames", [
    ("lenia_multikernel.json", {"injectAmount": 0.35, "muJitter": 0.4, "dt": 0.3}, 24),
    ("lenia_basic.json", {"growthMu": 0.16, "noiseAmount": 0.2}, 12),
])
def test_module_matches_scene_runner(driver, ctx, tmp_path, fname, bus, frames):
    """The module stepped frame by frame through encode(), with ParameterBus
    values that override the scene's own, against lenia_scene_step on the same
    scene with the same values set: field and colour layer identical."""
    import lenia

    path = os.path.join(SCENES, fname)
    scene, mods, W, H = _module(path)
    m = mods[0]
    name = m["name"]
    pj = tmp_path / "params.json"
    pj.write_text(json.dumps(m["params"]))
    bt = tmp_path / "bus.txt"
    bt.write_text("".join(f"{name}.{k} {v!r}\n" for k, v in bus.items()))
    st, lay = tmp_path / "state.bin", tmp_path / "rgba.bin"
    seed = int(scene.get("seed", 0))
    print(f"binding {fname}: {frames} frames, bus {bus}", flush=True)
    r = subprocess.run([driver, "run", str(pj), name, str(W), str(H), str(seed), str(frames), "1", str(bt),
                        str(st), str(lay)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with lenia.Scene(ctx, open(path).read(), 1) as sc:
        for k, v in bus.items():
            sc.set_param(f"{name}.{k}", v)
        sc.step(frames)
        want_state = sc.world(0).read()
        want_layer = sc.layer(0)
    got_state = np.fromfile(st, np.float32).reshape(want_state.shape)
    got_layer = np.fromfile(lay, np.float16).reshape(want_layer.shape)
    live = float(((want_state > 0) & (want_state < 1)).mean())
    print(f"binding {fname}: live fraction {live:.3f}", flush=True)
    assert live > 1e-3
    assert got_state.tobytes() == want_state.tobytes()
    assert got_layer.tobytes() == want_layer.tobytes()
