"""GPU seed extension (ga_extend_batch on gfx950) against the oracle: status,
score, full trace and the per-slice DP words of every kept slice."""
import numpy as np
import pytest

import oracle_ffi as o
import graphaligner_amd as ga
from workloads import extension_cases, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0)


def _run(ctx, preset, n, workdir, tweak=None, seed=7):
    gfa, names, seqs = workload(preset, workdir)
    og = o.Graph(gfa)
    idx = o.Index(og, 15, 20, 0.001)
    oparams = o.preset_params("vg")
    gparams = ga.params_preset("vg")
    if tweak:
        tweak(oparams, gparams)
    cases = extension_cases(og, idx, seqs, n, seed=seed)
    print(f"[gpu-extend] preset={preset} cases={len(cases)} running", flush=True)
    g = ga.Graph(gfa)
    got, ms = ga.extend_batch(ctx, g, gparams, [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases], summaries=True)
    bad = 0
    for (seq, node, off), (st, score, trace, rows) in zip(cases, got):
        want = o.extend(og, oparams, seq, node, off)
        same = want[0] == st and want[1] == score and np.array_equal(want[2], trace) and want[3] == rows
        bad += not same
    assert bad == 0, f"{bad} of {len(cases)} extensions differ"
    return got


@pytest.mark.parametrize("preset,n", [("mini", 200), ("mini_sv", 200), ("vg100k", 120)])
def test_gpu_extension_matches_oracle(ctx, preset, n, workdir):
    _run(ctx, preset, n, workdir)


def test_gpu_tangle_effort_matches_oracle(ctx, workdir):
    def tweak(op, gp):
        op.bandwidth = gp.bandwidth = 3
        op.maxCellsPerSlice = gp.maxCellsPerSlice = 60
    got = _run(ctx, "mini_sv", 100, workdir, tweak, seed=11)
    assert sum(r[9] for g in got for r in g[3]) > 0
