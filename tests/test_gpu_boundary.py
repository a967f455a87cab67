"""The reference's per-read boundary calls on gfx950 through the C ABI:
AlignOneWay with the caller's seeds (ga_align_one_way, src/GraphAlignerWrapper.h:55),
the AlignmentResult accessors, SelectAlignments and AddMappingQualities on the
result, AddGAFLine (ga_gaf_line) against the oracle; and the resident read set
rejected once another call replaced it (GA_ESTALE)."""
import numpy as np
import pytest

import oracle_ffi as o
import graphaligner_amd as ga
from workloads import edge_reads, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0)


def _text(res, read):
    """the result of one read in the layout of the oracle's align_one_way"""
    st = res.status(read)
    lines = ["S %d %d" % (st, res.seeds_extended(read) if st == 0 else 0)]
    if st == 0:
        for i in range(res.count(read)):
            it = res.item(read, i)
            lines.append("A %d %d %d %s %d %d %d" % (it["alignmentStart"], it["alignmentEnd"], it["alignmentScore"],
                                                    "%.17g" % it["alignmentXScore"], it["seedGoodness"],
                                                    it["mappingQuality"], it["traceLength"]))
            lines.append(" ".join("%d,%d,%d,%d,%d,%d" % (n, off, sp, sw, sc[0], gc[0]) for n, off, sp, sw, sc, gc in res.trace(read, i)))
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("wl,preset,n", [("mini_sv", "vg", 40), ("mini_sv", "dbg", 40), ("vg100k", "vg", 60)])
def test_gpu_align_one_way_matches_oracle(ctx, workdir, wl, preset, n):
    gfa, names, seqs = workload(wl, workdir)
    sd = o.PRESET_SEEDING[preset]
    og = o.Graph(gfa)
    oi = o.Index(og, sd["k"], sd["w"], sd["discard"])
    op = o.preset_params(preset)
    gp = ga.params_preset(preset)
    g = ga.Graph(gfa)
    en, es = edge_reads(names, seqs)
    rnames = list(names[:n]) + en
    reads = list(seqs[:n]) + es
    seeds = []
    for s in reads:
        x = oi.seeds(s, sd["density"], order=True)
        seeds.append(np.zeros((0, 10), np.int64) if x is None else x)
    print(f"[gpu-boundary] align_one_way {wl} {preset} reads={len(reads)}", flush=True)
    res = ga.align_one_way(ctx, g, gp, reads, seeds)
    assert len(res) == len(reads)
    for i, s in enumerate(reads):
        assert _text(res, i) == o.align_one_way(og, op, s, seeds[i]), (i, rnames[i])
    # MAPQ, selection and output order, then the GAF lines
    ot, ostat, _ = o.align_reads(og, oi, op, sd["density"], rnames, reads, 8)
    lines = []
    for i, s in enumerate(reads):
        if res.status(i) != 0:
            continue
        res.add_mapping_qualities(i)
        res.select(i, gp)
        assert _text(res, i) == o.align_one_way(og, op, s, seeds[i], select=True, mapq=True)
        # back to the default MAPQ of the non-multiseed path for the GAF comparison
    res2 = ga.align_one_way(ctx, g, gp, reads, seeds)
    for i, s in enumerate(reads):
        if res2.status(i) != 0 or len(seeds[i]) == 0:
            continue
        res2.select(i, gp)
        for k in range(res2.count(i)):
            lines.append(res2.gaf_line(g, i, k, rnames[i], len(s)) + "\n")
    assert "".join(lines) == ot


def test_gpu_resident_reads_go_stale(ctx, workdir):
    gfa, names, seqs = workload("mini", workdir)
    gp = ga.params_preset("vg")
    g = ga.Graph(gfa)
    gi = ga.Index(g, gp)
    print("[gpu-boundary] stale resident reads", flush=True)
    ga.upload_reads(ctx, seqs[:20])
    a, sa, _ = ga.align_resident(ctx, g, gi, gp, 0, names[:20])
    ga.align_reads(ctx, g, gi, gp, names[20:30], seqs[20:30])
    with pytest.raises(RuntimeError, match="GA_ESTALE"):
        ga.align_resident(ctx, g, gi, gp, 0, names[:5])
    ga.seeds(ctx, g, gi, seqs[:3], 10.0)
    with pytest.raises(RuntimeError, match="GA_ESTALE"):
        ga.align_resident(ctx, g, gi, gp, 0, names[:5])
    ga.upload_reads(ctx, seqs[:20])
    b, sb, _ = ga.align_resident(ctx, g, gi, gp, 0, names[:20])
    assert a == b and list(sa) == list(sb)
