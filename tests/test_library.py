"""The C-ABI library on the CPU: it loads, exports every entry point include/*.h declares, and
refuses (never falls back) without a GPU. No compute call touches a device here."""

import glob
import os
import re

import pytest

from cells_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(cells_\w+)\s*\(", text))
    return sorted(names)


def test_every_declared_entry_point_is_exported(hip_lib):
    names = declared()
    assert len(names) >= 15
    for name in names:
        assert hasattr(hip_lib, name), name
    assert set(names) == set(lib.EXPORTS)


def test_the_product_never_imports_the_oracle():
    for path in glob.glob(os.path.join(ROOT, "cells_amd", "**", "*.*"), recursive=True):
        if path.endswith((".py", ".hip", ".cpp", ".h")):
            text = open(path).read()
            assert "import oracle" not in text and "from oracle" not in text, path
            assert "cells_ref" not in text, path


def test_no_gpu_means_an_error_not_a_fallback():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from cells_amd.ensemble import Ensemble
    from cells_amd.world import build, config0

    with pytest.raises(RuntimeError):
        Ensemble(build(config0()))


def test_spatial_worlds_without_a_gpu_are_an_error():
    """cells_field_create checks its arguments, then refuses with CELLS_ENODEV (no fallback)."""
    import ctypes

    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from cells_amd import lib
    from cells_amd.field import build_field, small_field

    fp = build_field(small_field())
    w, g = lib.world_struct(fp.cells), lib.field_grid_struct(fp)
    h = ctypes.c_void_p()
    L = lib.load()
    assert L.cells_field_create(ctypes.byref(w), ctypes.byref(g), 0, 0, ctypes.byref(h)) == lib.CELLS_EINVAL
    assert L.cells_field_create(ctypes.byref(w), ctypes.byref(g), 50, 0, ctypes.byref(h)) == lib.CELLS_ENODEV
    fp.cell_node[1] = 0  # a cell on the held ring
    g = lib.field_grid_struct(fp)
    assert L.cells_field_create(ctypes.byref(w), ctypes.byref(g), 50, 0, ctypes.byref(h)) == lib.CELLS_EINVAL


def test_the_packed_fp32_switch_reaches_the_device_pass():
    """The build's guard: the probe kernel has packed FP32 instructions without the switch and
    none with it (gfx950 assembly of code built here)."""
    from cells_amd import build

    assert build.packed_fp32_count([]) > 0
    assert build.packed_fp32_count(build.NO_PACKED_FP32) == 0
