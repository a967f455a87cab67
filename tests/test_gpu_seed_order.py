"""Seed expansion and OrderSeeds on the GPU (ga_seeds with order 2,
ga_params_t.deviceSeedOrder) against the oracle's seeds and GAF: seeds field by
field in order (the GPU restates std::sort's tie order), whole-read GAF text and
statuses with the GPU ordering in the rounds and device-loop forms."""
import numpy as np
import pytest

import oracle_ffi as o
import graphaligner_amd as ga
from workloads import edge_reads, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0)


@pytest.mark.parametrize("wl,preset,n", [("mini_sv", "vg", 200), ("vg5m", "vg", 1500), ("mini_dbg", "dbg", 300)])
def test_gpu_ordered_seeds_match_oracle(ctx, workdir, wl, preset, n):
    gfa, names, seqs = workload(wl, workdir)
    sd = o.PRESET_SEEDING[preset]
    og = o.Graph(gfa)
    oi = o.Index(og, sd["k"], sd["w"], sd["discard"])
    gp = ga.params_preset(preset)
    g = ga.Graph(gfa)
    gi = ga.Index(g, gp, ctx=ctx)
    base = list(seqs[:n])
    reads = base + edge_reads(names, seqs)[1]
    print(f"[gpu-seed-order] {wl} {preset} reads={len(reads)}", flush=True)
    dev = ga.seeds(ctx, g, gi, reads, sd["density"], order=2)
    host = ga.seeds(ctx, g, gi, reads, sd["density"], order=1)
    compared = 0
    for i, (s, a, b) in enumerate(zip(reads, dev, host)):
        assert a.shape == b.shape and np.array_equal(a, b), i
        if i < 120 or i >= len(base):
            want = oi.seeds(s, sd["density"], True)
            if want is None:
                assert len(a) == 0
            else:
                assert want.shape == a.shape and (want == a).all(), i
        compared += len(a)
    assert compared > 1000


@pytest.mark.parametrize("loop", [0, 1])
@pytest.mark.parametrize("wl,preset", [("mini_sv", "vg"), ("mini_dbg", "dbg")])
def test_gpu_gaf_with_device_seed_order(ctx, workdir, wl, preset, loop):
    gfa, names, seqs = workload(wl, workdir)
    sd = o.PRESET_SEEDING[preset]
    og = o.Graph(gfa)
    oi = o.Index(og, sd["k"], sd["w"], sd["discard"])
    gp = ga.params_preset(preset)
    gp.deviceSeedOrder = 1
    gp.deviceSeedLoop = loop
    g = ga.Graph(gfa)
    gi = ga.Index(g, gp)
    en, es = edge_reads(names, seqs)
    rn, rs = list(names[:150]) + en, list(seqs[:150]) + es
    print(f"[gpu-seed-order] gaf {wl} {preset} loop={loop} reads={len(rs)}", flush=True)
    text, status, _ = ga.align_reads(ctx, g, gi, gp, rn, rs)
    want, wstat, _ = o.align_reads(og, oi, o.preset_params(preset), sd["density"], rn, rs, 8)
    assert list(status) == list(wstat)
    assert text == want
