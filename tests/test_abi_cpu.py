"""CPU checks of the C-ABI library: it loads, exports every symbol the public
header declares, and its struct mirrors have the header's layout.  No compute
calls are made here."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lenia_mi355x.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lenia_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from lenia import _abi

    L = _abi.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/lenia_mi355x.h but not exported"
    assert sorted(_abi.EXPORTS) == names


def test_struct_layouts():
    from lenia import _abi

    # lenia_kernel_desc: 6 scalars + betas[4] + 3 scalars, all 4 bytes
    assert C.sizeof(_abi.KernelDesc) == 4 * 13
    # lenia_world_desc: 11 leading scalars, 4 kernels, 7 init fields, 6 tuning fields
    assert C.sizeof(_abi.WorldDesc) == 4 * 11 + 4 * 52 + 4 * 7 + 4 * 6 + 4 * 2


def test_version_and_last_error():
    from lenia import _abi

    L = _abi.lib()
    assert b"gfx950" in L.lenia_version()
    assert isinstance(L.lenia_last_error(), bytes)


def test_null_arguments_are_rejected_without_a_device():
    from lenia import _abi

    L = _abi.lib()
    assert L.lenia_ctx_create(0, None) == _abi.ERR_INVALID_ARG
    assert L.lenia_world_create(None, None, None) == _abi.ERR_INVALID_ARG
    assert L.lenia_world_step(None, 1) == _abi.ERR_INVALID_ARG
    assert L.lenia_world_read_state(None, None, 0) == _abi.ERR_INVALID_ARG
    assert L.lenia_ctx_sync(None) == _abi.ERR_INVALID_ARG
    assert b"NULL" in L.lenia_last_error()


def test_world_desc_marshalling():
    from lenia import _abi
    from lenia_cases import config1

    d = _abi.world_desc(config1())
    assert (d.width, d.height, d.nKernels) == (1024, 512, 3)
    assert d.kernels[2].nBetas == 3
    assert abs(d.kernels[2].betas[2] - 0.25) < 1e-7
    assert (d.initX0, d.initY0, d.initW, d.initH) == (384, 128, 256, 256)


def test_sweep_worlds_are_valid():
    from lenia_cases import SWEEP_SIDES, sweep_worlds

    ws = sweep_worlds(256)
    assert len(ws) == 256
    for w in ws:
        assert w["width"] in SWEEP_SIDES and w["height"] in SWEEP_SIDES
        assert 1 <= len(w["kernels"]) <= 3
        R = w["kernels"][0]["radius"]
        assert 2 * R + 1 <= min(w["width"], w["height"])
        assert w["dt"] in (0.05, 0.1, 0.2)


def _build_c_caller(tmp_path):
    import shutil
    import subprocess

    from lenia import _abi

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lib = _abi.LIB_PATH
    exe = str(tmp_path / "abi_caller")
    subprocess.run([cc, "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "abi_caller.c"), "-o", exe,
                    lib, "-Wl,-rpath," + os.path.dirname(lib)], check=True)
    return exe


def test_c_caller_against_the_header(tmp_path):
    """A plain C program compiled against include/lenia_mi355x.h: the header's
    sizeof / offsetof equal the Python mirror's; every NULL-argument call
    returns LENIA_ERR_INVALID_ARG and leaves a message for lenia_last_error;
    lenia_scene_describe resolves the reference's multikernel preset through
    the C ABI (no GPU call is made)."""
    import subprocess

    from lenia import _abi

    exe = _build_c_caller(tmp_path)
    scene = os.path.join(ROOT, "tests", "golden", "scenes", "lenia_multikernel.json")
    out = subprocess.run([exe, scene], check=True, capture_output=True, text=True).stdout
    mirrors = {"lenia_kernel_desc": _abi.KernelDesc, "lenia_world_desc": _abi.WorldDesc,
               "lenia_param_grads": _abi.ParamGrads, "lenia_colormap": _abi.ColorMap,
               "lenia_scene_info": _abi.SceneInfo}
    seen = {k: set() for k in mirrors}
    for line in out.splitlines():
        f = line.split()
        if f[0] == "size":
            assert C.sizeof(mirrors[f[1]]) == int(f[2]), line
        elif f[0] == "offset":
            assert getattr(mirrors[f[1]], f[2]).offset == int(f[3]), line
            seen[f[1]].add(f[2])
    for name, m in mirrors.items():
        assert seen[name] == {fld for fld, _ in m._fields_}, name
    rcs = [ln for ln in out.splitlines() if ln.startswith("rc ")]
    assert len(rcs) == 11
    for ln in rcs:
        assert f" {_abi.ERR_INVALID_ARG} " in ln, ln
        assert "NULL" in ln or "null" in ln.lower(), ln
    assert "gfx950" in out
    desc = [ln for ln in out.splitlines() if ln.startswith("describe")][0].split()
    # Presets/lenia_multikernel.json: simWidth/simHeight 1024x512, FFT, growth mode 1,
    # three kernels, radius 24, dt 0.35
    assert desc[2] == "0" and desc[4] == "lenia0"
    assert desc[6] == "1024x512" and desc[8] == "1" and desc[10] == "1" and desc[12] == "3"
    assert desc[14] == "24" and desc[16] == "0.350"
    kern = [ln.split() for ln in out.splitlines() if ln.startswith("kernel ")]
    assert [k[3] for k in kern] == ["24.000", "13.200", "7.200"]


def test_plan_cache_set_get_save_load(tmp_path):
    """the radix-plan cache is host state: entries are validated against their
    length, removed with plan 0, saved as text and loaded all-or-nothing"""
    import lenia
    from lenia import _abi

    lenia.plan_cache_clear()
    assert lenia.plan_cache_get(lenia.ROWS, 1024) == 0
    lenia.plan_cache_set(lenia.ROWS, 1024, 0x88F)   # 16, 8, 8
    lenia.plan_cache_set(lenia.COLS, 512, 0x48F)    # 16, 8, 4
    lenia.plan_cache_set(lenia.COLS, 96, 0x384)     # 4, 8, 3
    assert lenia.plan_cache_get(lenia.ROWS, 1024) == 0x88F
    assert lenia.plan_cache_get(lenia.COLS, 1024) == 0
    for axis, n, plan in ((0, 1024, 0x88), (1, 96, 0x88), (2, 64, 0x88), (0, 100, 0x55)):
        with pytest.raises(lenia.LeniaError) as ei:
            lenia.plan_cache_set(axis, n, plan)
        assert ei.value.code == _abi.ERR_INVALID_ARG
    path = tmp_path / "plans.txt"
    lenia.plan_cache_save(path)
    text = path.read_text()
    assert "rows 1024 0x88f" in text and "cols 512 0x48f" in text and "cols 96 0x384" in text
    lenia.plan_cache_clear()
    assert lenia.plan_cache_get(lenia.COLS, 512) == 0
    lenia.plan_cache_load(path)
    assert lenia.plan_cache_get(lenia.ROWS, 1024) == 0x88F
    assert lenia.plan_cache_get(lenia.COLS, 96) == 0x384
    lenia.plan_cache_set(lenia.ROWS, 1024, 0)
    assert lenia.plan_cache_get(lenia.ROWS, 1024) == 0
    bad = tmp_path / "bad.txt"
    bad.write_text("# comment\nrows 256 0xff\ncols 256 0x3\n")  # 3 does not divide 256
    with pytest.raises(lenia.LeniaError):
        lenia.plan_cache_load(bad)
    assert lenia.plan_cache_get(lenia.ROWS, 256) == 0  # nothing of the bad file was taken
    lenia.plan_cache_clear()
