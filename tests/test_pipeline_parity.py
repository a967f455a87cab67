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
@pytest.mark.parametrize("wl,preset", [("mini", "vg"), ("mini", "dbg"), ("mini_sv", "vg"), ("mini_sv", "dbg"), ("mini_dbg", "dbg"),
                                       ("mini_dbg", "vg")])
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


def _homopolymer_workload(workdir):
    """A small graph whose sequence holds single-base runs of 12-60 bp, and reads
    with runs of every length around k and the window (w - k + 1), lowercase and
    N-interrupted runs: the cases of iterateKmers' repeat rule."""
    import random
    rng = random.Random(11)
    parts = []
    for r in range(40):
        parts.append("".join(rng.choice("ACGT") for _ in range(rng.randrange(30, 90))))
        if r % 5 == 0:
            parts.append(rng.choice("ACGT") * rng.randrange(12, 60))
    genome = "".join(parts)
    path = os.path.join(workdir, "homopolymer.gfa")
    chunks = [genome[i:i + 700] for i in range(0, len(genome), 700)]
    with open(path, "w") as f:
        f.write("H\tVN:Z:1.0\n")
        for i, c in enumerate(chunks):
            f.write(f"S\t{i + 1}\t{c}\n")
        for i in range(len(chunks) - 1):
            f.write(f"L\t{i + 1}\t+\t{i + 2}\t+\t0M\n")
    reads = []
    for run in list(range(10, 40)) + [45, 50, 57, 64, 80, 101, 150]:
        for base in "ACGT":
            left = "".join(rng.choice("ACGT") for _ in range(40))
            right = "".join(rng.choice("ACGT") for _ in range(40))
            reads.append(left + base * run + right)
    for run in (30, 60, 90):
        reads.append("GATTACA" * 5 + "a" * run + "CCGTA" * 6)
        reads.append("TTGCA" * 6 + "C" * (run // 2) + "N" + "C" * (run // 2) + "AGGT" * 8)
        reads.append("ACGT" * 5 + "G" * run + "n" + "g" * run + "TACG" * 5)
    # pieces of the genome itself, runs included
    for _ in range(40):
        a = rng.randrange(0, len(genome) - 400)
        reads.append(genome[a:a + rng.randrange(100, 400)])
    return path, reads


@pytest.mark.parametrize("preset", ["vg", "dbg"])
def test_seeds_of_single_base_runs(workdir, preset):
    gfa, reads = _homopolymer_workload(workdir)
    sd, og, oi, pg, pi = _pair(gfa, preset)
    nonempty = 0
    for s in reads:
        for order in (False, True):
            a = oi.seeds(s, sd["density"], order, cap_per_base=512)
            b = pi.seeds(s, sd["density"], order)
            assert a.shape == b.shape
            assert (a == b).all()
            nonempty += len(a) > 0
    assert nonempty > len(reads)


@pytest.mark.parametrize("mode", sorted(P.MODES))
def test_gaf_identical_wide_fans(workdir, mode):
    """Slices across 48-way fans outgrow the first-attempt band-queue capacities:
    the overflowing extensions are rerun with larger workspaces."""
    gfa, names, seqs = W.fan_workload(workdir)
    sd, og, oi, pg, pi = _pair(gfa, "vg")
    params = O.preset_params("vg")
    ot, ostat, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 4)
    pt, pstat, _ = P.align_reads(pi, params, sd["density"], names, seqs, **P.MODES[mode])
    assert (ostat == 0).sum() >= len(seqs) // 2
    assert list(pstat) == list(ostat)
    assert pt == ot


@pytest.mark.parametrize("mode", ["rounds_spec", "device_loop", "batched"])
def test_gaf_identical_dense_seeds(workdir, mode):
    """Minimizer length 8 with the window equal to k and no k-mers discarded:
    268,498 seeds on 50 reads, about 20,000 extensions and hundreds of kept
    alignments per read, so every seed's coverage test runs against long lists of
    alignments, in the host rounds and in the host builds of the device seed loop
    and of the batched persistent form."""
    gfa, names, seqs = W.workload("mini_sv", workdir)
    en, es = W.edge_reads(names, seqs)
    names = list(names) + en
    seqs = list(seqs) + es
    density = O.PRESET_SEEDING["vg"]["density"]
    oi = O.Index(O.Graph(gfa), 8, 8, 0.0)
    pi = P.Index(P.Graph(gfa), 8, 8, 0.0)
    params = O.preset_params("vg")
    ot, ostat, _ = O.align_reads(oi.graph, oi, params, density, names, seqs, 8)
    pt, pstat, stats = P.align_reads(pi, params, density, names, seqs, **P.MODES[mode])
    assert stats["seeds"] > 200000 and stats["extensions"] > 10000
    assert len(ot.splitlines()) > 200
    assert list(pstat) == list(ostat)
    assert pt == ot


@pytest.mark.parametrize("mode", ["rounds_spec", "device_loop"])
@pytest.mark.parametrize("wl,preset", [("mini_sv", "vg"), ("mini_dbg", "dbg")])
def test_gaf_identical_cigar_match_mismatch_merge(workdir, wl, preset, mode):
    """--cigar-match-mismatch (src/Aligner.cpp:564): M runs in the cigar, matches
    and mismatches still counted one by one for the match count, NM, dv and id."""
    gfa, names, seqs = W.workload(wl, workdir)
    sd, og, oi, pg, pi = _pair(gfa, preset)
    en, es = W.edge_reads(names, seqs)
    names = list(names) + en
    seqs = list(seqs) + es
    params = O.preset_params(preset)
    plain, _, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 4)
    params.cigarMatchMismatchMerge = 1
    ot, ostat, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 4)
    pt, pstat, _ = P.align_reads(pi, params, sd["density"], names, seqs, **P.MODES[mode])
    assert list(pstat) == list(ostat)
    assert pt == ot
    # the option changes only the cigar: M in place of = and X
    assert "M" in ot and ot != plain
    assert [l.split("\tcg:Z:")[0] for l in ot.splitlines()] == [l.split("\tcg:Z:")[0] for l in plain.splitlines()]
