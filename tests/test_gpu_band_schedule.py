"""Per-slice bandwidth schedules on the GPU (ga_extend_batch_scheduled on gfx950)
against the oracle under the same schedules: status, score, whole trace and the
per-slice DP words of every kept slice; a schedule of equal entries against
ga_extend_batch."""
import random

import numpy as np
import pytest

import oracle_ffi as o
import graphaligner_amd as ga
from test_band_schedule import random_schedule
from workloads import extension_cases, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0)


@pytest.mark.parametrize("wl,preset,cap,n", [("mini_sv", "vg", 0, 160), ("vg100k", "vg", 0, 80), ("mini_dbg", "dbg", 400, 120)])
def test_gpu_schedules_match_oracle(ctx, workdir, wl, preset, cap, n):
    gfa, names, seqs = workload(wl, workdir)
    og = o.Graph(gfa)
    sd = o.PRESET_SEEDING[preset]
    idx = o.Index(og, sd["k"], sd["w"], sd["discard"])
    op = o.preset_params(preset)
    gp = ga.params_preset(preset)
    if cap:
        op.maxCellsPerSlice = gp.maxCellsPerSlice = cap
    cases = extension_cases(og, idx, seqs, n, seed=31, density=sd["density"])
    rng = random.Random(41)
    # every fourth extension keeps the fixed bandwidth (empty schedule)
    sched = [random_schedule(rng, len(c[0])) if i % 4 else [] for i, c in enumerate(cases)]
    print(f"[gpu-schedule] {wl} {preset} cases={len(cases)} running", flush=True)
    g = ga.Graph(gfa)
    got, ms = ga.extend_batch(ctx, g, gp, [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases],
                              summaries=True, schedules=sched)
    bad = 0
    for (seq, node, off), s, (st, score, trace, rows) in zip(cases, sched, got):
        want = o.extend(og, op, seq, node, off, schedule=s)
        bad += not (want[0] == st and want[1] == score and np.array_equal(want[2], trace) and want[3] == rows)
    assert bad == 0, f"{bad} of {len(cases)} extensions differ"


def test_gpu_constant_schedule_equals_fixed_band(ctx, workdir):
    gfa, names, seqs = workload("mini_sv", workdir)
    og = o.Graph(gfa)
    idx = o.Index(og, 15, 20, 0.001)
    gp = ga.params_preset("vg")
    cases = extension_cases(og, idx, seqs, 60, seed=37)
    g = ga.Graph(gfa)
    args = ([c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases])
    fixed, _ = ga.extend_batch(ctx, g, gp, *args, summaries=True)
    const, _ = ga.extend_batch(ctx, g, gp, *args, summaries=True, schedules=[[gp.bandwidth]] * len(cases))
    for a, b in zip(fixed, const):
        assert a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2]) and a[3] == b[3]


def test_gpu_schedule_rejects_bad_bandwidths(ctx, workdir):
    gfa, names, seqs = workload("mini", workdir)
    og = o.Graph(gfa)
    idx = o.Index(og, 15, 20, 0.001)
    cases = extension_cases(og, idx, seqs, 2, seed=3)
    g = ga.Graph(gfa)
    gp = ga.params_preset("vg")
    for bad in ([-1], [5, (1 << 20) + 1]):
        with pytest.raises(RuntimeError, match="GA_EBADARG"):
            ga.extend_batch(ctx, g, gp, [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases],
                            schedules=[bad] * len(cases))
