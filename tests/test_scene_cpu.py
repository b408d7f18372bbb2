"""Scene JSON -> world descriptor mapping (lenia_scene_describe, no GPU) and
the oracle's colour-map, stamp and activity functions, pinned by closed forms.

The expected descriptors are what LeniaModule::setup
(Modules/FieldModules/Lenia.cpp:208-356) resolves for the reference's own
presets (tests/golden/scenes/, copied verbatim from Presets/ and sweeps/):
simWidth/simHeight 0 -> the scene resolution, convMode "fft" only for powers
of two, per-kernel radius = float(round(radius)) * radiusScale, the growth
fallbacks (0.15 / 0.017 / dt 0.1, or the organism's), and the palette of
ColorMapPass::configure on top of Lenia's defaults.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = os.path.join(HERE, "golden", "scenes")


def scene_text(name, edit=None):
    with open(os.path.join(SCENES, name)) as f:
        j = json.load(f)
    if edit:
        edit(j)
    return json.dumps(j)


def describe(text, module=0):
    import lenia

    return lenia.describe_scene(text, module)


def f32(x):
    return float(np.float32(x))


def test_multikernel_preset():
    d, cm, name = describe(scene_text("lenia_multikernel.json"))
    import lenia

    assert name == "lenia0"
    assert (d["width"], d["height"]) == (1024, 512)  # simWidth/simHeight, not the 1280x720 resolution
    assert d["convMode"] == lenia.CONV_FFT and d["radius"] == 24
    assert d["dt"] == f32(0.35) and d["growthMode"] == 1
    assert d["injectAmount"] == f32(0.2) and d["injectCount"] == 4
    assert d["initMode"] == lenia.INIT_NOISE
    assert d["initCoverage"] == f32(0.42) and d["initScale"] == 26.0
    ks = d["kernels"]
    assert len(ks) == 3
    # radius * radiusScale in FP32, ring shells with kernelShellSigma
    assert [k["radius"] for k in ks] == [24.0, f32(np.float32(24) * np.float32(0.55)),
                                         f32(np.float32(24) * np.float32(0.3))]
    assert [k["betas"] for k in ks] == [[1.0], [1.0, f32(0.6)], [1.0]]
    assert [k["mu"] for k in ks] == [f32(0.15), f32(0.30), f32(0.22)]
    assert [k["weight"] for k in ks] == [1.0, f32(0.6), f32(0.4)]
    assert all(k["shape"] == lenia.SHAPE_RING and k["shellSigma"] == f32(0.15) for k in ks)
    assert cm["d"] == [f32(0.6), f32(0.42), f32(0.3)] and cm["channel"] == 0


def test_basic_preset_direct():
    import lenia

    d, cm, _ = describe(scene_text("lenia_basic.json"))
    assert (d["width"], d["height"]) == (768, 768)
    assert d["convMode"] == lenia.CONV_DIRECT and d["radius"] == 13
    k = d["kernels"][0]
    assert len(d["kernels"]) == 1 and k["shape"] == lenia.SHAPE_SHELL and k["radius"] == 13.0
    assert (k["shellMu"], k["shellSigma"]) == (0.5, f32(0.15))
    assert (k["mu"], k["sigma"]) == (f32(0.15), f32(0.017))
    assert d["injectAmount"] == f32(0.22) and d["injectCount"] == 4 and d["growthMode"] == 1


def test_fft_needs_powers_of_two_and_kernels_need_fft():
    import lenia

    # 1280x720 is not a power of two: the reference falls back to direct,
    # taking radiusScale / mu / sigma of kernels[0] (Lenia.cpp:245-269)
    def edit(j):
        j["modules"][0]["params"].update(simWidth=0, simHeight=0)

    d, _, _ = describe(scene_text("lenia_multikernel.json", edit))
    assert d["convMode"] == lenia.CONV_DIRECT and (d["width"], d["height"]) == (1280, 720)
    assert d["radius"] == 24 and d["kernels"][0]["mu"] == f32(0.15)

    def edit2(j):
        p = j["modules"][0]["params"]
        p.update(convMode="direct")
        p["kernels"][0].update(radiusScale=0.5, mu=0.2, sigma=0.03)
        p.pop("growthMu")
        p.pop("growthSigma")

    d, _, _ = describe(scene_text("lenia_multikernel.json", edit2))
    assert d["convMode"] == lenia.CONV_DIRECT and d["radius"] == 12
    assert (d["kernels"][0]["mu"], d["kernels"][0]["sigma"]) == (f32(0.2), f32(0.03))


def test_fft_without_kernels_uses_growth_params():
    import lenia

    def edit(j):
        p = j["modules"][0]["params"]
        p.update(convMode="fft", simWidth=512, simHeight=256)

    d, _, _ = describe(scene_text("lenia_basic.json", edit))
    assert d["convMode"] == lenia.CONV_FFT and len(d["kernels"]) == 1
    k = d["kernels"][0]
    assert k["shape"] == lenia.SHAPE_RING and k["radius"] == 13.0 and k["betas"] == [1.0]
    assert (k["mu"], k["sigma"], k["weight"]) == (f32(0.15), f32(0.017), 1.0)


def test_more_than_four_kernels_truncate():
    def edit(j):
        p = j["modules"][0]["params"]
        p["kernels"] = p["kernels"] + p["kernels"]

    d, _, _ = describe(scene_text("lenia_multikernel.json", edit))
    assert len(d["kernels"]) == 4


def test_organism_scene_params_and_stamp():
    import lenia

    def edit(j):
        j["modules"][0]["params"]["organism"] = os.path.join(SCENES, "orbium.json")

    d, _, _ = describe(scene_text("lenia_orbium.json", edit))
    # R=13, mu=0.15, sigma=0.015, T=10 -> dt 0.1 (Lenia.cpp:315-320)
    assert d["radius"] == 13 and d["initMode"] == lenia.INIT_STAMP
    assert d["dt"] == f32(0.1)
    assert (d["kernels"][0]["mu"], d["kernels"][0]["sigma"]) == (f32(0.15), f32(0.015))

    # a missing organism file: noise init, module defaults
    def edit2(j):
        j["modules"][0]["params"]["organism"] = "/nonexistent/orbium.json"

    d, _, _ = describe(scene_text("lenia_orbium.json", edit2))
    assert d["initMode"] == lenia.INIT_NOISE and d["dt"] == f32(0.1) and d["radius"] == 13
    assert d["kernels"][0]["sigma"] == f32(0.017)


def test_mono_palette_and_freeze_probe():
    d, cm, _ = describe(scene_text("freeze_probe.json"))
    assert (d["width"], d["height"]) == (384, 216) and d["radius"] == 9
    assert cm["a"] == cm["b"] == cm["c"] == cm["d"] == [0.5, 0.5, 0.5]


def test_scene_errors():
    import lenia

    with pytest.raises(lenia.LeniaError):
        describe("{not json")
    with pytest.raises(lenia.LeniaError):
        describe(json.dumps({"modules": [{"name": "x"}]}))  # module without "type"
    with pytest.raises(lenia.LeniaError) as e:
        describe(json.dumps({"modules": [{"type": "Slime"}]}))
    assert e.value.code == -2
    with pytest.raises(lenia.LeniaError):
        describe(json.dumps({"modules": [{"type": "Lenia", "params": {"radius": "big"}}]}))
    # disabled modules do not count
    with pytest.raises(lenia.LeniaError):
        describe(json.dumps({"modules": [{"type": "Lenia", "enabled": False}]}))


# ---------------------------------------------------------------- oracle pins

PAL = dict(a=[.5, .5, .5], b=[.5, .5, .5], c=[1, 1, 1], d=[.6, .45, .3])


def test_colormap_closed_forms():
    f = np.array([[0.0, 0.25, 1.0, 1.5], [-0.5, 0.5, 0.75, 0.1]])
    out = O.colormap(f, PAL)
    t = np.clip(f, 0, 1)
    for ch in range(3):
        want = np.maximum(0.5 + 0.5 * np.cos(2 * np.pi * (t + PAL["d"][ch])), 0)
        np.testing.assert_allclose(out[..., ch], want, atol=1e-15)
    np.testing.assert_array_equal(out[..., 3], t)
    # the "mono" palette is the ramp 0.5 - 0.5 cos(pi t): black at 0, white at 1
    mono = dict(a=[.5] * 3, b=[.5] * 3, c=[.5] * 3, d=[.5] * 3)
    o = O.colormap(np.array([[0.0, 1.0, 0.5]]), mono)
    np.testing.assert_allclose(o[0, :, 0], [0.0, 1.0, 0.5], atol=1e-15)
    # bias / scale / exposure and the channel rule of an R32F texel
    o = O.colormap(np.array([[0.2]]), dict(PAL, inputScale=2.0, inputBias=0.1, exposure=3.0))
    assert o[0, 0, 3] == pytest.approx(0.5)
    assert o[0, 0, 0] == pytest.approx(3 * (0.5 + 0.5 * np.cos(2 * np.pi * 1.1)))
    assert O.colormap(np.array([[0.7]]), dict(PAL, channel=1))[0, 0, 3] == 0.0
    assert O.colormap(np.array([[0.7]]), dict(PAL, channel=3))[0, 0, 3] == 1.0


def test_colormap_bilinear():
    rng = np.random.default_rng(3)
    f = rng.random((6, 8))
    # identity c=1, d=0 palette is not linear, so read the interpolated value
    # back through alpha with inputScale 1: alpha = clamp(v)
    # downsample by 2: pixel centres fall half-way between texel pairs
    o = O.colormap(f, PAL, 4, 3)[..., 3]
    want = 0.25 * (f[0::2, 0::2] + f[0::2, 1::2] + f[1::2, 0::2] + f[1::2, 1::2])
    np.testing.assert_allclose(o, want, atol=1e-15)
    # upsample by 2 with wrap: output x=0 sits at texel -0.25 -> 0.25 f[W-1] + 0.75 f[0]
    o = O.colormap(f, PAL, 16, 12)[..., 3]
    row = 0.75 * f[0] + 0.25 * f[-1]  # y: pixel 0 at texel -0.25 as well
    assert o[0, 0] == pytest.approx(0.75 * row[0] + 0.25 * row[-1])
    # a constant field stays constant at any size
    o = O.colormap(np.full((5, 7), 0.3), PAL, 13, 9)[..., 3]
    np.testing.assert_allclose(o, 0.3, atol=1e-15)


def test_luma_stats():
    px = np.zeros((4, 4, 4))
    px[..., :3] = 0.5
    m, sd = O.luma_stats(px)
    assert m == pytest.approx(0.5) and sd == pytest.approx(0.0, abs=1e-12)
    px[:2, :, 1] = 1.0  # half the pixels gain 0.7152 * 0.5 in luma
    m, sd = O.luma_stats(px)
    assert m == pytest.approx(0.5 + 0.7152 * 0.25) and sd == pytest.approx(0.7152 * 0.25)


def dihedral(a):
    out = []
    for f in (a, a[:, ::-1]):
        for k in range(4):
            out.append(np.rot90(f, k))
    return out


def test_stamp_geometry():
    org = json.load(open(os.path.join(SCENES, "orbium.json")))["cells"]
    org = np.array(org, np.float64)
    seen = set()
    for seed in range(40):
        out = O.init_stamp(96, 80, org, 1, 1, seed)
        # one stamp: the organism, whole, in one of 8 orientations somewhere
        # on the torus (roll it away from the border first)
        ys, xs = np.nonzero(out)
        out = np.roll(out, (40 - int(ys[0]), 48 - int(xs[0])), (0, 1))
        ys, xs = np.nonzero(out)
        box = out[ys.min():ys.max() + 1, xs.min():xs.max() + 1]
        match = [i for i, t in enumerate(dihedral(org)) if t.shape == box.shape and np.array_equal(t, box)]
        assert match, seed
        seen.add(match[0])
    assert len(seen) >= 6  # the 8 orientations all occur
    # rotation off: always the organism as stored
    out = O.init_stamp(96, 80, org, 1, 0, 12345)
    ys, xs = np.nonzero(out)
    out = np.roll(out, (40 - int(ys[0]), 48 - int(xs[0])), (0, 1))
    ys, xs = np.nonzero(out)
    assert np.array_equal(out[ys.min():ys.max() + 1, xs.min():xs.max() + 1], org)
    # several stamps: the maximum of the copies
    many = O.init_stamp(96, 80, org, 6, 1, 7)
    assert many.max() == org.max() and (many > 0).sum() >= (out > 0).sum()
