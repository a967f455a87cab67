"""The device extension algorithm built for the host (tests/emu) against the oracle:
traces, scores and per-slice DP words. "wave": lane helpers as loops over 64
lanes; "thread": the one-extension-per-thread helpers the gfx950 kernel runs."""
import numpy as np
import pytest

import emu_ffi as e
import oracle_ffi as o
from workloads import extension_cases, workload


@pytest.mark.parametrize("mode", ["wave", "thread"])
@pytest.mark.parametrize("preset,n", [("mini", 80), ("mini_sv", 80), ("vg100k", 24)])
def test_emulated_device_logic_matches_oracle(preset, n, mode, workdir):
    gfa, names, seqs = workload(preset, workdir)
    og = o.Graph(gfa)
    eg = e.Graph(gfa, mode)
    idx = o.Index(og, 15, 20, 0.001)
    params = o.preset_params("vg")
    for seq, node, off in extension_cases(og, idx, seqs, n):
        a = o.extend(og, params, seq, node, off)
        b = e.extend(eg, params, seq, node, off)
        assert a[0] == b[0] and a[1] == b[1]
        assert np.array_equal(a[2], b[2])
        assert a[3] == b[3]


@pytest.mark.parametrize("mode", ["wave", "thread"])
def test_emulated_tangle_effort_and_small_band(mode, workdir):
    gfa, names, seqs = workload("mini_sv", workdir)
    og = o.Graph(gfa)
    eg = e.Graph(gfa, mode)
    idx = o.Index(og, 15, 20, 0.001)
    params = o.preset_params("vg")
    params.bandwidth = 3
    params.maxCellsPerSlice = 60
    notvalid = 0
    for seq, node, off in extension_cases(og, idx, seqs, 40, seed=11):
        a = o.extend(og, params, seq, node, off)
        b = e.extend(eg, params, seq, node, off)
        assert a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2]) and a[3] == b[3]
        notvalid += sum(r[9] for r in a[3])
    assert notvalid > 0


@pytest.mark.parametrize("mode", ["wave", "thread"])
def test_emulated_dbg_tangle_effort(mode, workdir):
    """-x dbg (b = 5) on the cyclic de Bruijn graph with the tangle-effort cap
    lowered so that capped slices (scoresNotValid) occur."""
    gfa, names, seqs = workload("mini_dbg", workdir)
    og = o.Graph(gfa)
    eg = e.Graph(gfa, mode)
    sd = o.PRESET_SEEDING["dbg"]
    idx = o.Index(og, sd["k"], sd["w"], sd["discard"])
    params = o.preset_params("dbg")
    params.maxCellsPerSlice = 400
    notvalid = 0
    for seq, node, off in extension_cases(og, idx, seqs, 40, seed=13, density=sd["density"]):
        a = o.extend(og, params, seq, node, off)
        b = e.extend(eg, params, seq, node, off)
        assert a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2]) and a[3] == b[3]
        notvalid += sum(r[9] for r in a[3])
    assert notvalid > 0
