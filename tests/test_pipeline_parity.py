"""The product's host pipeline (index, seeding, seed loop, selection, GAF) over a
host build of the gfx950 device code, compared with the oracle: minimizer index,
seeds (MinimizerSeeder::getSeeds + OrderSeeds) and whole GAF output byte for byte
(src/Aligner.cpp:390-606), for the vg and dbg presets."""
import os

import numpy as np
import pytest

import oracle_ffi as O
import pipe_ffi as P
import workloads as W
from test_oracle_golden import README_LINE, GOLD, read_fa


def _pair(gfa, preset):
    sd = O.PRESET_SEEDING[preset]
    og = O.Graph(gfa)
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    pg = P.Graph(gfa)
    pi = P.Index(pg, sd["k"], sd["w"], sd["discard"])
    return sd, og, oi, pg, pi


def test_readme_known_answer():
    gfa = os.path.join(GOLD, "readme_graph.gfa")
    sd, og, oi, pg, pi = _pair(gfa, "vg")
    name, seq = read_fa(os.path.join(GOLD, "readme_read.fa"))
    text, status, _ = P.align_reads(pi, O.preset_params("vg"), sd["density"], [name], [seq])
    assert list(status) == [0]
    assert text == README_LINE + "\n"


@pytest.mark.parametrize("wl", ["mini", "mini_sv"])
@pytest.mark.parametrize("preset", ["vg", "dbg"])
def test_index_and_seeds(workdir, wl, preset):
    gfa, names, seqs = W.workload(wl, workdir)
    sd, og, oi, pg, pi = _pair(gfa, preset)
    assert pi.max_count == oi.max_count
    assert pi.num_kmers == lib_num_kmers(oi)
    en, es = W.edge_reads(names, seqs)
    for s in list(seqs[:12]) + es:
        for order in (False, True):
            a = oi.seeds(s, sd["density"], order)
            b = pi.seeds(s, sd["density"], order)
            assert a.shape == b.shape
            assert (a == b).all()


def lib_num_kmers(oracle_index):
    return O.lib().orc_index_num_kmers(oracle_index.h)


@pytest.mark.parametrize("mode", sorted(P.MODES))
@pytest.mark.parametrize("wl", ["mini", "mini_sv"])
@pytest.mark.parametrize("preset", ["vg", "dbg"])
def test_gaf_identical(workdir, wl, preset, mode):
    gfa, names, seqs = W.workload(wl, workdir)
    sd, og, oi, pg, pi = _pair(gfa, preset)
    en, es = W.edge_reads(names, seqs)
    names = list(names) + en
    seqs = list(seqs) + es
    params = O.preset_params(preset)
    ot, ostat, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 4)
    pt, pstat, st = P.align_reads(pi, params, sd["density"], names, seqs, **P.MODES[mode])
    assert list(pstat) == list(ostat)
    assert set(ostat.tolist()) >= {0, 1, 2}
    assert pt == ot
    assert st["extensions"] >= int((pstat == 0).sum())


@pytest.mark.parametrize("mode", sorted(P.MODES))
def test_gaf_identical_vg100k_sample(workdir, mode):
    gfa, names, seqs = W.workload("vg100k", workdir)
    names, seqs = names[:120], seqs[:120]
    sd, og, oi, pg, pi = _pair(gfa, "vg")
    params = O.preset_params("vg")
    ot, ostat, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 8)
    pt, pstat, _ = P.align_reads(pi, params, sd["density"], names, seqs, **P.MODES[mode])
    assert list(pstat) == list(ostat)
    assert pt == ot


@pytest.mark.parametrize("mode", sorted(P.MODES))
@pytest.mark.parametrize("wl", ["mini", "mini_sv"])
@pytest.mark.parametrize("preset", ["vg", "dbg"])
def test_gaf_identical_all_seeds_cluster2(workdir, wl, preset, mode):
    """--try-all-seeds (no sloppy shortcuts) with --seeds-clustersize 2."""
    gfa, names, seqs = W.workload(wl, workdir)
    sd, og, oi, pg, pi = _pair(gfa, preset)
    en, es = W.edge_reads(names, seqs)
    names = list(names) + en
    seqs = list(seqs) + es
    params = O.preset_params(preset)
    params.sloppy = 0
    params.minClusterSize = 2
    ot, ostat, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 4)
    pt, pstat, _ = P.align_reads(pi, params, sd["density"], names, seqs, **P.MODES[mode])
    assert list(pstat) == list(ostat)
    assert pt == ot
