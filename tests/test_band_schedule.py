"""Per-slice bandwidth schedule: the hook through which getXdropSlices picks each
slice's bandwidth (src/GraphAlignerBitvectorBanded.h:611), kept in
DPSlice.bandwidth (:552) and read back by the traceback's quit scores
(src/GraphAlignerBitvectorCommon.h:633,659,684).

The oracle's schedule is pinned by
* the reference's own debug-only invariants (assertSliceCorrectness on every
  column, checkNodeBoundaryCorrectness on every slice with that slice's quit
  score minimum + bandwidth, src/GraphAlignerBitvectorCommon.h:1062-1076,
  src/GraphAlignerBitvectorBanded.h:169-227) holding under random schedules;
* a schedule of equal entries reproducing the fixed bandwidth exactly, and a
  schedule that only changes slices past the last one kept changing nothing;
* a schedule of unbounded bandwidths reproducing the fixed unbounded band, whose
  slice minima equal a textbook edit-distance DP (test_oracle_pinning.py).
The device algorithm built for the host is compared with the oracle under the
same schedules (the GPU comparison is in test_gpu_band_schedule.py).
"""
import random

import numpy as np
import pytest

import emu_ffi as E
import oracle_ffi as O
import workloads as W


def random_schedule(rng, seq_len, lo=0, hi=16):
    n = seq_len // 64 + 1
    kind = rng.randrange(4)
    if kind == 0:  # one value
        return [rng.randrange(lo, hi)]
    if kind == 1:  # wide near the seed, then narrow (a ramp)
        k = rng.randrange(1, n + 1)
        return [rng.randrange(8, hi + 8)] * k + [rng.randrange(lo, 8)]
    if kind == 2:  # short random prefix, the last entry holds
        return [rng.randrange(lo, hi) for _ in range(rng.randrange(1, 6))]
    return [rng.randrange(lo, hi) for _ in range(n)]  # every slice its own


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2]) and a[3] == b[3]


@pytest.mark.parametrize("wl,preset,cap", [("mini_sv", "vg", 0), ("mini_sv", "vg", 200), ("mini_dbg", "dbg", 0),
                                           ("mini_dbg", "dbg", 400)])
def test_reference_invariants_hold_under_schedules(workdir, wl, preset, cap):
    gfa, names, seqs = W.workload(wl, workdir)
    og = O.Graph(gfa, ndebug="checked")
    sd = O.PRESET_SEEDING[preset]
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    p = O.preset_params(preset)
    if cap:
        p.maxCellsPerSlice = cap
    L = og.L
    rng = random.Random(5)
    run0, fail0 = L.orc_extra_checks_run(), L.orc_extra_check_failures()
    kept = 0
    for c in W.extension_cases(og, oi, seqs, 40, seed=21, density=sd["density"]):
        r = O.extend(og, p, *c, schedule=random_schedule(rng, len(c[0])))
        kept += len(r[3])
    runs, fails = L.orc_extra_checks_run() - run0, L.orc_extra_check_failures() - fail0
    assert kept > 100
    assert runs > 100_000
    assert fails == 0


@pytest.mark.parametrize("wl,preset", [("mini_sv", "vg"), ("mini_dbg", "dbg")])
def test_constant_schedule_is_the_fixed_band(workdir, wl, preset):
    gfa, names, seqs = W.workload(wl, workdir)
    og = O.Graph(gfa)
    sd = O.PRESET_SEEDING[preset]
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    p = O.preset_params(preset)
    b = p.bandwidth
    differ = 0
    for c in W.extension_cases(og, oi, seqs, 30, seed=4, density=sd["density"]):
        fixed = O.extend(og, p, *c)
        n = len(c[0]) // 64 + 1
        assert same(fixed, O.extend(og, p, *c, schedule=[b]))
        assert same(fixed, O.extend(og, p, *c, schedule=[b] * n))
        # entries past the last kept slice are never read
        assert same(fixed, O.extend(og, p, *c, schedule=[b] * len(fixed[3]) + [0, 1, 2]))
        # a different band changes something in most extensions
        other = O.extend(og, p, *c, schedule=[b + 7])
        differ += not same(fixed, other)
    assert differ > 0


def test_unbounded_schedule_is_the_unbounded_band(workdir):
    """The unbounded band (with X-drop off) is pinned against a textbook
    edit-distance DP over the graph in test_oracle_pinning.py; an unbounded
    schedule must give the same tables."""
    gfa, names, seqs = W.workload("mini_sv", workdir)
    og = O.Graph(gfa)
    oi = O.Index(og, 15, 20, 0.001)
    p = O.preset_params("vg")
    p.bandwidth = 1 << 20
    p.xdrop = 1 << 30
    for c in W.extension_cases(og, oi, seqs, 12, seed=9):
        c = (c[0][:700], c[1], c[2])
        assert same(O.extend(og, p, *c), O.extend(og, p, *c, schedule=[1 << 20, 1 << 20]))


@pytest.mark.parametrize("mode", ["wave", "thread"])
@pytest.mark.parametrize("wl,preset,cap", [("mini_sv", "vg", 0), ("mini_dbg", "dbg", 400)])
def test_emulated_device_logic_matches_oracle_under_schedules(workdir, mode, wl, preset, cap):
    gfa, names, seqs = W.workload(wl, workdir)
    og = O.Graph(gfa)
    eg = E.Graph(gfa, mode)
    sd = O.PRESET_SEEDING[preset]
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    p = O.preset_params(preset)
    if cap:
        p.maxCellsPerSlice = cap
    rng = random.Random(17)
    for c in W.extension_cases(og, oi, seqs, 40, seed=23, density=sd["density"]):
        sch = random_schedule(rng, len(c[0]))
        a = O.extend(og, p, *c, schedule=sch)
        b = E.extend(eg, p, *c, schedule=sch)
        assert same(a, b), sch
