"""The C ABI library loads without a GPU and exports every declared entry point."""
import ctypes
import os
import re

import pytest

from graphaligner_amd import _native as N

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "graphaligner_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:ga_status_t|void|int64_t|int32_t|const\s+\w+\s*\*)\s*(ga_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(N.LIB_PATH)
    names = declared()
    assert len(names) >= 10
    for name in names:
        assert hasattr(L, name), name


def test_presets():
    import graphaligner_amd as ga
    vg = ga.params_preset("vg")
    assert (vg.bandwidth, vg.minimizerLength, vg.minimizerWindowSize, vg.minimizerSeedDensity, vg.seedExtendDensity) == (10, 15, 20, 10.0, -1.0)
    dbg = ga.params_preset("dbg")
    assert (dbg.bandwidth, dbg.maxCellsPerSlice, dbg.minimizerLength, dbg.minimizerWindowSize) == (5, 10000, 19, 30)
    with pytest.raises(RuntimeError):
        ga.params_preset("nope")


def test_no_gpu_fails_loudly():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    import graphaligner_amd as ga
    with pytest.raises(RuntimeError, match="GA_ENODEVICE"):
        ga.Context(0)
