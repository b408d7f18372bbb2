"""Scene runner, colour layer, organism stamps and parameter sweeps on the GPU.

* colour map (lenia_colormap_field) vs the FP64 oracle of colorMapField /
  colorMapFieldScaled: within one half-float ulp of the oracle value (the
  layer is RGBA16F) plus, when resampling, the stated bound of the FP32
  texel positions.  The reference samples with the GPU's texture filter,
  whose weight precision is not specified: its exact bits are unpinned;
* stamp init vs the oracle's leniaStampInit: bit-exact (same FP32 geometry);
* a scene stepped by the runner equals a world built from the same scene
  descriptor and stepped frame by frame, bit for bit, including the per-frame
  parameter, freeze, reseed and visibility rules of SceneRunner::step;
* every variant of a sweep equals its own world stepped alone, bit for bit;
  the GPU activity metric agrees with the oracle's luma statistics of the
  oracle colour layer of that state.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = os.path.join(HERE, "golden", "scenes")
PAL = dict(a=[.5, .5, .5], b=[.5, .5, .5], c=[1, 1, 1], d=[.6, .45, .3])


@pytest.fixture(scope="module")
def ctx():
    import lenia

    c = lenia.Context(0)
    yield c
    c.close()


def scene_text(name, edit=None):
    with open(os.path.join(SCENES, name)) as f:
        j = json.load(f)
    if edit:
        edit(j)
    return json.dumps(j)


def half_ulp(want):
    w = np.abs(np.asarray(want, np.float64)).astype(np.float16)
    return np.spacing(np.maximum(w, np.float16(2.0 ** -14))).astype(np.float64)


def check_layer(got, want, what, extra=1e-6):
    """`extra`: bound on the FP32-vs-FP64 difference before rounding to half"""
    got = np.asarray(got, np.float64)
    err = np.abs(got - want)
    tol = half_ulp(want) + extra
    bad = int((err > tol).sum())
    print(f"{what}: max err {err.max():.2e}, {bad} of {err.size} beyond one half ulp")
    assert bad == 0


@pytest.mark.parametrize("dst", [(128, 96), (160, 120), (64, 48), (200, 75)])
def test_colormap_field_matches_oracle(ctx, dst):
    import lenia

    rng = np.random.default_rng(5)
    field = rng.uniform(-0.2, 1.2, (96, 128)).astype(np.float32)
    params = dict(PAL, inputScale=1.3, inputBias=-0.05, exposure=1.5)
    dev = torch.tensor(field, device="cuda")
    dw, dh = dst
    out = torch.zeros((dh, dw, 4), dtype=torch.float16, device="cuda")
    torch.cuda.synchronize()
    lenia.colormap_field(ctx, dev.data_ptr(), 128, 96, params, out.data_ptr(), dw, dh)
    ctx.sync()
    got = out.cpu().numpy()
    want = O.colormap(field.astype(np.float64), params, dw, dh)
    # resampled: the FP32 texel position u*W - 0.5 is off by up to ~W * 2^-23
    # texels (1.5e-5 at W = 128), moving v by that times the neighbour step
    # (< 1.4 here), then scaled by the palette slope 2 pi b c inputScale
    # exposure (< 6.2): below 2e-4
    extra = 1e-6 if (dw, dh) == (128, 96) else 2e-4
    check_layer(got, want, f"colormap 128x96 -> {dw}x{dh}", extra)


def orbium():
    return np.array(json.load(open(os.path.join(SCENES, "orbium.json")))["cells"], np.float64)


@pytest.mark.parametrize("count,rotate,seed", [(6, True, 7777), (1, False, 42), (32, True, 9)])
def test_stamp_init_bit_exact(ctx, count, rotate, seed):
    import lenia
    from lenia_cases import DIRECT, gauss_shell

    org = orbium()
    W, H = 200, 120
    # dt = 0: the frame's step leaves the initial field as it is
    d = dict(width=W, height=H, convMode=DIRECT, growthMode=0, dt=0.0, radius=5,
             kernels=[gauss_shell(5)], initMode=lenia.INIT_STAMP)
    with lenia.World(ctx, d) as w:
        w.set_organism(org, count, rotate)
        w.reset(seed)
        w.frame(0, 1)
        got = w.read()
    want = O.init_stamp(W, H, org, count, rotate, O.module_seed(seed)).astype(np.float32)
    print(f"stamps={count} rotate={rotate}: {(want > 0).sum()} cells set")
    assert np.array_equal(got, want)


def world_of_scene(ctx, text, module=0):
    import lenia

    d, cm, name = lenia.describe_scene(text, module)
    return lenia.World(ctx, d), cm


def test_multikernel_scene_matches_world(ctx):
    import lenia

    text = scene_text("lenia_multikernel.json")
    with lenia.Scene(ctx, text) as sc:
        sc.step(3)
        info = sc.info()
        assert info["frameIndex"] == 3 and (info["width"], info["height"]) == (1280, 720)
        got = sc.world(0).read()
        layer = sc.layer(0)
    w, cm = world_of_scene(ctx, text)
    with w:
        w.reset(7777)
        for f in range(3):
            w.frame(f, 1)
        want = w.read()
    assert np.array_equal(got, want)
    print(f"1024x512 K=3 after 3 frames: mean {got.mean():.4f}")
    # the 1280x720 layer is the bilinear colour map of the 1024x512 field
    # resampled 1024x512 -> 1280x720: FP32 texel positions within 1024 * 2^-23
    # texels, neighbour steps < 1, palette slope pi
    check_layer(layer, O.colormap(got.astype(np.float64), cm, 1280, 720), "multikernel layer", 5e-4)


def test_multikernel_preset_300_frames_acceptance(ctx):
    """The reference's acceptance run of the multikernel preset
    (docs/specs/phase9_lenia_fft.md:163-164): Presets/lenia_multikernel.json
    for 300 frames, twice -- every call succeeds, the final frame is not one
    colour, and both runs give identical bytes (state and colour layer).  Then
    frame 300 from the run's own state against the FP64 oracle (growth mode 1,
    inject blobs at that frame's seed), within 1e-5."""
    import hashlib

    import lenia
    from lenia_cases import live_cells

    text = scene_text("lenia_multikernel.json")
    runs = []
    for _ in range(2):
        with lenia.Scene(ctx, text) as sc:
            sc.step(300)
            state = sc.world(0).read()
            layer = sc.layer(0)
            runs.append((state, layer))
            if len(runs) == 1:
                prev = state
                sc.step(1)
                nxt = sc.world(0).read()
    (s0, l0), (s1, l1) = runs
    md5 = [hashlib.md5(l.tobytes()).hexdigest() for l in (l0, l1)]
    rgb = l0[..., :3].astype(np.float64).reshape(-1, 3)
    colours = len(np.unique(l0[..., :3].reshape(-1, 3), axis=0))
    luma = 0.2126 * rgb[:, 0] + 0.7152 * rgb[:, 1] + 0.0722 * rgb[:, 2]
    print(f"multikernel 300 frames: layer md5 {md5[0]} / {md5[1]}; {colours} distinct colours, "
          f"luma mean {luma.mean():.3f} sd {luma.std():.3f}; {live_cells(s0)} live cells")
    assert s0.tobytes() == s1.tobytes() and md5[0] == md5[1]
    assert colours > 1 and luma.std() > 1e-3
    d, _, _ = lenia.describe_scene(text)
    m = O.module_seed(7777)
    want = O.step(d, prev.astype(np.float64), use_fft=True, seed=O.step_seed(m, 300, 0), frame=300)
    mx = float(np.abs(nxt.astype(np.float64) - want).max())
    print(f"multikernel frame 300 from the GPU state vs FP64 oracle: max-abs {mx:.2e}")
    assert mx <= 1e-5


def test_scene_frame_rules(ctx):
    """per-frame parameters, freeze, reseed and visibility (SceneRunner.cpp:146-215)"""
    import lenia

    text = scene_text("freeze_probe.json")
    seed = 4242
    d, cm, _ = lenia.describe_scene(text)

    def fresh_from(sd, frames):
        with lenia.World(ctx, d) as w:
            w.reset(sd)
            for f in frames:
                w.frame(f, 1)
            return w.read()

    with lenia.Scene(ctx, text) as sc:
        wv = sc.world(0)
        sc.step(2)
        s2 = wv.read()
        assert np.array_equal(s2, fresh_from(seed, [0, 1]))
        # freeze: no encode while high; frames still count
        sc.set_param("lenia0.freeze", 1.0)
        sc.step(3)
        assert np.array_equal(wv.read(), s2) and sc.info()["frameIndex"] == 5
        sc.set_param("lenia0.freeze", 0.0)
        # reseed rising edge at frame 5: reset(seed ^ frame*2654435761), init, step
        sc.set_param("lenia0.reseed", 1.0)
        sc.step(1)
        s6 = wv.read()
        assert np.array_equal(s6, fresh_from(seed ^ ((5 * 2654435761) & 0xFFFFFFFF), [5]))
        sc.step(1)  # still high: no second reseed
        assert np.array_equal(wv.read(), fresh_from(seed ^ ((5 * 2654435761) & 0xFFFFFFFF), [5, 6]))
        sc.set_param("lenia0.reseed", 0.0)
        # hidden: not stepped; shown again at frame 9: reborn from the frame seed
        sc.set_param("lenia0.opacity", 0.0)
        s7 = wv.read()
        sc.step(2)
        assert np.array_equal(wv.read(), s7)
        sc.set_param("lenia0.opacity", 1.0)
        sc.step(1)
        assert np.array_equal(wv.read(), fresh_from(seed ^ ((9 * 2654435761) & 0xFFFFFFFF), [9]))
        # dt = 0 through the parameter bus: growth mode 1 leaves the field as is
        before = wv.read()
        sc.set_param("lenia0.dt", 0.0)
        sc.step(1)
        assert np.array_equal(wv.read(), before)
        layer = sc.layer(0)
    check_layer(layer, O.colormap(before.astype(np.float64), cm), "freeze probe layer (mono)")


def test_orbium_scene_stamps(ctx):
    import lenia

    path = os.path.join(SCENES, "orbium.json")

    def edit(j):
        j["modules"][0]["params"]["organism"] = path

    text = scene_text("lenia_orbium.json", edit)
    d, cm, _ = lenia.describe_scene(text)
    with lenia.Scene(ctx, text) as sc:
        sc.step(2)
        got = sc.world(0).read()
    with lenia.World(ctx, d) as w:
        w.set_organism(orbium(), 8, True)
        w.reset(7777)
        w.frame(0, 1)
        w.frame(1, 1)
        want = w.read()
    assert np.array_equal(got, want)
    # and the init of that scene is the oracle's stamp field
    init = O.init_stamp(d["width"], d["height"], orbium(), 8, True, O.module_seed(7777))
    with lenia.World(ctx, dict(d, dt=0.0)) as w:
        w.set_organism(orbium(), 8, True)
        w.reset(7777)
        w.frame(0, 1)
        assert np.array_equal(w.read(), init.astype(np.float32))


def check_sweep(ctx, scene, sweep, frames, variants_expected):
    import lenia

    with lenia.Sweep(ctx, scene, sweep) as sw:
        assert sw.n == variants_expected
        sw.run(frames)
        mean, sd, alive = sw.activity()
        rep = sw.report()
        states = [sw.world(v).read() for v in range(sw.n)]
    assert rep["variantCount"] == variants_expected and rep["frames"] == frames
    for v, entry in enumerate(rep["variants"]):
        assert entry["id"] == f"v{v:04d}"
        sj = json.loads(scene)
        sj["seed"] = entry["seed"]
        for target, value in entry["overrides"].items():
            mod, key = target.split(".", 1)
            sj["modules"][0]["params"][key] = value
        d, cm, _ = lenia.describe_scene(json.dumps(sj))
        with lenia.World(ctx, d) as w:
            w.reset(entry["seed"])
            for f in range(frames):
                w.frame(f, 1)
            want = w.read()
        assert np.array_equal(states[v], want), entry
        info = json.loads(scene)
        W, H = info["resolution"]
        layer = O.colormap(want.astype(np.float64), cm, W, H).astype(np.float16)
        m, s = O.luma_stats(layer)
        assert abs(mean[v] - m) < 1e-6 and abs(sd[v] - s) < 1e-6, (v, mean[v], m, sd[v], s)
        assert entry["activity"]["alive"] == bool(alive[v])
    return rep, mean, sd, alive


def test_sweep_per_axis_direct(ctx):
    scene = scene_text("lenia_basic.json")
    sweep = open(os.path.join(SCENES, "lenia_mu_sigma.json")).read()
    rep, mean, sd, alive = check_sweep(ctx, scene, sweep, 3, 12)
    # per-axis: seeds outermost, then each axis' values in order (main.cpp:253-265)
    seeds = [e["seed"] for e in rep["variants"]]
    assert seeds == [7777] * 6 + [1234] * 6
    targets = [list(e["overrides"].items())[0] for e in rep["variants"][:6]]
    assert [t[0] for t in targets] == ["lenia0.growthMu"] * 3 + ["lenia0.growthSigma"] * 3
    np.testing.assert_allclose([t[1] for t in targets], [0.13, 0.15, 0.17, 0.014, 0.018, 0.022])
    print("activity:", np.round(mean, 4), np.round(sd, 4), alive.astype(int))


def test_sweep_cartesian_fft(ctx):
    scene = scene_text("lenia_multikernel.json")
    sweep = json.dumps({"mode": "cartesian", "seeds": [11, 12],
                        "axes": [{"target": "lenia0.dt", "values": [0.2, 0.35]},
                                 {"target": "lenia0.injectAmount", "range": [0.0, 0.2], "steps": 2}]})
    rep, mean, sd, alive = check_sweep(ctx, scene, sweep, 4, 8)
    # odometer order, last axis fastest
    ov = [tuple(e["overrides"].values()) for e in rep["variants"][:4]]
    assert ov == [(0.2, 0.0), (0.2, 0.2), (0.35, 0.0), (0.35, 0.2)]


def test_sweep_errors(ctx):
    import lenia

    scene = scene_text("lenia_basic.json")
    with pytest.raises(lenia.LeniaError):
        lenia.Sweep(ctx, scene, json.dumps({"mode": "spiral"}))
    with pytest.raises(lenia.LeniaError):
        lenia.Sweep(ctx, scene, json.dumps({"axes": [{"target": "growthMu", "values": [1]}]}))
    with pytest.raises(lenia.LeniaError):
        lenia.Sweep(ctx, scene, json.dumps({"axes": [{"target": "lenia0.dt", "range": [0, 1], "steps": 0}]}))
    # more than 500 variants are refused (main.cpp:684-687)
    big = {"axes": [{"target": "lenia0.dt", "range": [0.1, 0.2], "steps": 30},
                    {"target": "lenia0.growthMu", "range": [0.1, 0.2], "steps": 20}]}
    with pytest.raises(lenia.LeniaError):
        lenia.Sweep(ctx, scene, json.dumps(big))
    # an override of an unknown module is that variant's error, not the sweep's
    with lenia.Sweep(ctx, scene, json.dumps({"axes": [{"target": "other.dt", "values": [0.1]}]})) as sw:
        sw.run(1)
        rep = sw.report()
        assert "unknown module" in rep["variants"][0]["error"]
        assert rep["variants"][0]["activity"]["alive"] is False


def test_npy_state_roundtrip(ctx, tmp_path):
    """.npy snapshots (NpyWriter.cpp's NumPy format): numpy reads what the
    library writes and the library loads what numpy writes, bit for bit; bad
    files are refused with the world unchanged"""
    import lenia
    from lenia_cases import config4

    d = config4()
    with lenia.World(ctx, d) as w:
        w.reset(31)
        w.frame(0, 2)
        A = w.read()
        p = tmp_path / "state.npy"
        w.save_npy(p)
        raw = p.read_bytes()
        assert raw[:6] == b"\x93NUMPY" and (10 + int.from_bytes(raw[8:10], "little")) % 64 == 0
        B = np.load(p)
        assert B.dtype == np.float32 and B.shape == A.shape and np.array_equal(A, B)
        rng = np.random.default_rng(2)
        C_ = rng.random(A.shape, dtype=np.float32)
        q = tmp_path / "in.npy"
        np.save(q, C_)
        w.load_npy(q)
        assert np.array_equal(w.read(), C_)
        # stepping from a loaded state equals stepping from the same write
        w.step(3)
        after = w.read()
    with lenia.World(ctx, d) as w2:
        w2.reset(31)
        w2.write(C_)
        w2.step(3)
        assert np.array_equal(w2.read(), after)
        for bad in (C_.astype(np.float64), C_[:, :-1], np.asfortranarray(C_.T).T.copy(order="F"),
                    C_.astype(">f4")):
            r = tmp_path / "bad.npy"
            np.save(r, bad)
            before = w2.read()
            with pytest.raises(lenia.LeniaError):
                w2.load_npy(r)
            assert np.array_equal(w2.read(), before)
        r = tmp_path / "short.npy"
        np.save(r, C_)
        r.write_bytes(r.read_bytes()[:-8])
        with pytest.raises(lenia.LeniaError):
            w2.load_npy(r)
