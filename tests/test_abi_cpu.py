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
    assert C.sizeof(_abi.WorldDesc) == 4 * 11 + 4 * 52 + 4 * 7 + 4 * 6


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
