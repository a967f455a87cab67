"""Whole-read alignment on gfx950 through the C ABI (ga_index_build, ga_seeds,
ga_align_reads) against the oracle: minimizer index, seeds, per-read status and
the GAF text byte for byte (src/Aligner.cpp:390-606), vg and dbg presets."""
import os

import pytest

import oracle_ffi as o
import graphaligner_amd as ga
from test_oracle_golden import GOLD, README_LINE, read_fa
from workloads import edge_reads, fan_workload, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0)


def _both(gfa, preset):
    sd = o.PRESET_SEEDING[preset]
    og = o.Graph(gfa)
    oi = o.Index(og, sd["k"], sd["w"], sd["discard"])
    gp = ga.params_preset(preset)
    g = ga.Graph(gfa)
    gi = ga.Index(g, gp)
    assert gi.max_count == oi.max_count
    return sd, og, oi, gp, g, gi


def test_gpu_readme_known_answer(ctx):
    sd, og, oi, gp, g, gi = _both(os.path.join(GOLD, "readme_graph.gfa"), "vg")
    name, seq = read_fa(os.path.join(GOLD, "readme_read.fa"))
    print("[gpu-align] readme", flush=True)
    text, status, _ = ga.align_reads(ctx, g, gi, gp, [name], [seq])
    assert list(status) == [0]
    assert text == README_LINE + "\n"


@pytest.mark.parametrize("wl", ["mini", "mini_sv"])
@pytest.mark.parametrize("preset", ["vg", "dbg"])
def test_gpu_seeds_match_oracle(ctx, workdir, wl, preset):
    gfa, names, seqs = workload(wl, workdir)
    sd, og, oi, gp, g, gi = _both(gfa, preset)
    en, es = edge_reads(names, seqs)
    reads = list(seqs[:16]) + es
    print(f"[gpu-align] seeds {wl} {preset} reads={len(reads)}", flush=True)
    got = ga.seeds(ctx, g, gi, reads, sd["density"], order=True)
    for s, b in zip(reads, got):
        a = oi.seeds(s, sd["density"], True)
        assert a.shape == b.shape and (a == b).all()


@pytest.mark.parametrize("k", [5, 8, 15, 31])
def test_gpu_seeds_match_oracle_k_w(ctx, workdir, k):
    """Seeds at minimizer lengths beyond the presets' (5 to 31, the longest the
    index takes) with the window equal to k and no k-mers discarded, so that every
    k-mer is a minimizer and a run of one base longer than the window is reported
    at every position (src/MinimizerSeeder.cpp:105-189 through getSeeds); at
    k = 31 the lookup kernel's k-mer window reaches the start of its staged halo.
    (Wider windows leave this small graph's maxCount at 0 and no seeds.) The
    seeds expanded and ordered on the GPU equal the host's."""
    w = k
    gfa, names, seqs = workload("mini_sv", workdir)
    sd = o.PRESET_SEEDING["vg"]
    og = o.Graph(gfa)
    oi = o.Index(og, k, w, 0.0)
    gp = ga.params_preset("vg")
    gp.minimizerLength, gp.minimizerWindowSize = k, w
    gp.minimizerDiscardMostNumerousFraction = 0.0
    g = ga.Graph(gfa)
    gi = ga.Index(g, gp)
    assert gi.max_count == oi.max_count
    en, es = edge_reads(names, seqs)
    s0 = seqs[0]
    runs = s0[:300] + "A" * 70 + s0[300:600] + "C" * 33 + s0[600:900] + "G" * 31 + s0[900:1200] + "T" * 2 + s0[1200:1500]
    reads = list(seqs[:12]) + es + [runs]
    print(f"[gpu-align] seeds mini_sv k={k} w={w} reads={len(reads)}", flush=True)
    got = ga.seeds(ctx, g, gi, reads, sd["density"], order=True)
    assert len(got) == len(reads)
    nonempty = 0
    for s, b in zip(reads, got):
        a = oi.seeds(s, sd["density"], True)
        if a is None:
            # OrderSeeds fired a reference assertion: the read gets no seeds
            assert len(b) == 0
            continue
        assert a.shape == b.shape and (a == b).all()
        nonempty += len(a) > 0
    assert nonempty >= 15
    # expansion and OrderSeeds on the GPU (order 2) give the same seeds
    dev = ga.seeds(ctx, g, gi, reads, sd["density"], order=2)
    assert all(x.shape == y.shape and (x == y).all() for x, y in zip(got, dev))


# execution modes of the seed loop: host-driven rounds one seed at a time, rounds
# with speculative extension of later seeds, the whole loop on the device
MODES = {"rounds": (0, 1), "rounds_spec": (0, 64), "device_loop": (1, 1)}


def _set_mode(gp, mode):
    gp.deviceSeedLoop, gp.maxSpeculativeSeeds = MODES[mode]


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("wl", ["mini", "mini_sv"])
@pytest.mark.parametrize("preset", ["vg", "dbg"])
def test_gpu_gaf_matches_oracle(ctx, workdir, wl, preset, mode):
    gfa, names, seqs = workload(wl, workdir)
    sd, og, oi, gp, g, gi = _both(gfa, preset)
    _set_mode(gp, mode)
    en, es = edge_reads(names, seqs)
    names = list(names) + en
    seqs = list(seqs) + es
    ot, ostat, _ = o.align_reads(og, oi, o.preset_params(preset), sd["density"], names, seqs, 8)
    print(f"[gpu-align] gaf {wl} {preset} reads={len(seqs)}", flush=True)
    gt, gstat, st = ga.align_reads(ctx, g, gi, gp, names, seqs)
    assert list(gstat) == list(ostat)
    assert gt == ot


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("k", [8, 15, 31])
def test_gpu_gaf_matches_oracle_k_w(ctx, workdir, k, mode):
    """Whole reads at minimizer lengths beyond the presets' (window equal to k, no
    k-mers discarded): at k = 8 thousands of seeds per read and several kept
    alignments per read, at k = 31 reads left without seeds."""
    if k == 8 and mode == "device_loop":
        pytest.skip("the device seed loop runs a read's seeds one after another on one lane, each "
                    "tested against every kept alignment: at k = 8 (268,498 seeds, ~20,000 extensions "
                    "on 50 reads) that is minutes of serial work per read; its host build gives the "
                    "oracle's GAF on this input in test_pipeline_parity.py::test_gaf_identical_dense_seeds")
    gfa, names, seqs = workload("mini_sv", workdir)
    sd = o.PRESET_SEEDING["vg"]
    og = o.Graph(gfa)
    oi = o.Index(og, k, k, 0.0)
    gp = ga.params_preset("vg")
    gp.minimizerLength, gp.minimizerWindowSize = k, k
    gp.minimizerDiscardMostNumerousFraction = 0.0
    _set_mode(gp, mode)
    g = ga.Graph(gfa)
    gi = ga.Index(g, gp)
    assert gi.max_count == oi.max_count
    en, es = edge_reads(names, seqs)
    names = list(names) + en
    seqs = list(seqs) + es
    ot, ostat, _ = o.align_reads(og, oi, o.preset_params("vg"), sd["density"], names, seqs, 8)
    print(f"[gpu-align] gaf mini_sv k={k} {mode} reads={len(seqs)}", flush=True)
    gt, gstat, _ = ga.align_reads(ctx, g, gi, gp, names, seqs)
    assert list(gstat) == list(ostat)
    assert gt == ot
    assert len(ot.splitlines()) >= 37


@pytest.mark.parametrize("mode", ["rounds_spec", "device_loop"])
def test_gpu_gaf_cigar_match_mismatch_merge(ctx, workdir, mode):
    """--cigar-match-mismatch through ga_params_t (src/Aligner.cpp:564): M runs in
    the cigar, matches and mismatches still counted apart."""
    gfa, names, seqs = workload("mini_sv", workdir)
    sd, og, oi, gp, g, gi = _both(gfa, "vg")
    _set_mode(gp, mode)
    gp.cigarMatchMismatchMerge = 1
    op = o.preset_params("vg")
    op.cigarMatchMismatchMerge = 1
    ot, ostat, _ = o.align_reads(og, oi, op, sd["density"], names, seqs, 8)
    print(f"[gpu-align] gaf cigar merge {mode} reads={len(seqs)}", flush=True)
    gt, gstat, _ = ga.align_reads(ctx, g, gi, gp, names, seqs)
    assert list(gstat) == list(ostat)
    assert "M" in ot
    assert gt == ot


@pytest.mark.parametrize("mode", sorted(MODES))
def test_gpu_gaf_matches_oracle_vg100k(ctx, workdir, mode):
    gfa, names, seqs = workload("vg100k", workdir)
    names, seqs = names[:400], seqs[:400]
    sd, og, oi, gp, g, gi = _both(gfa, "vg")
    _set_mode(gp, mode)
    ot, ostat, _ = o.align_reads(og, oi, o.preset_params("vg"), sd["density"], names, seqs, 16)
    print(f"[gpu-align] gaf vg100k reads={len(seqs)}", flush=True)
    gt, gstat, st = ga.align_reads(ctx, g, gi, gp, names, seqs)
    assert list(gstat) == list(ostat)
    assert gt == ot


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("wl", ["mini", "mini_sv"])
@pytest.mark.parametrize("preset", ["vg", "dbg"])
def test_gpu_gaf_all_seeds_cluster2(ctx, workdir, wl, preset, mode):
    """--try-all-seeds (no sloppy shortcuts) with --seeds-clustersize 2."""
    gfa, names, seqs = workload(wl, workdir)
    sd, og, oi, gp, g, gi = _both(gfa, preset)
    en, es = edge_reads(names, seqs)
    names = list(names) + en
    seqs = list(seqs) + es
    op = o.preset_params(preset)
    op.sloppy = 0
    op.minClusterSize = 2
    gp.tryAllSeeds = 1
    gp.seedClusterMinSize = 2
    _set_mode(gp, mode)
    ot, ostat, _ = o.align_reads(og, oi, op, sd["density"], names, seqs, 8)
    print(f"[gpu-align] all seeds {wl} {preset} reads={len(seqs)}", flush=True)
    gt, gstat, st = ga.align_reads(ctx, g, gi, gp, names, seqs)
    assert list(gstat) == list(ostat)
    assert gt == ot


@pytest.mark.parametrize("preset", ["vg", "dbg"])
def test_gpu_seeds_of_single_base_runs(ctx, workdir, preset):
    """The block-parallel lookup's per-position repeat rule on single-base runs."""
    from test_pipeline_parity import _homopolymer_workload
    gfa, reads = _homopolymer_workload(workdir)
    sd, og, oi, gp, g, gi = _both(gfa, preset)
    print(f"[gpu-align] single-base-run seeds {preset} reads={len(reads)}", flush=True)
    got = ga.seeds(ctx, g, gi, reads, sd["density"], order=True)
    for s, b in zip(reads, got):
        a = oi.seeds(s, sd["density"], True, cap_per_base=512)
        assert a.shape == b.shape and (a == b).all()


@pytest.mark.parametrize("preset", ["vg", "dbg"])
def test_gpu_seeds_of_runs_across_chunks(ctx, workdir, preset):
    """The lookup stages each 256-position chunk of a read with the 32 positions
    before it: single-base runs that end just after a chunk starts, straddle a
    chunk boundary, or fill a whole k-mer window across it."""
    import random
    from test_pipeline_parity import _homopolymer_workload
    gfa, _ = _homopolymer_workload(workdir)
    sd, og, oi, gp, g, gi = _both(gfa, preset)
    rng = random.Random(23)
    reads = []
    for start in (200, 226, 240, 250, 255, 256, 257, 270, 480, 500, 511):
        for run in (14, 16, 20, 33, 40, 70):
            base = rng.choice("ACGT")
            left = "".join(rng.choice("ACGT") for _ in range(start))
            right = "".join(rng.choice("ACGT") for _ in range(rng.randrange(40, 300)))
            reads.append(left + base * run + right)
        reads.append("".join(rng.choice("ACGT") for _ in range(start)) + "T" * 20 + "N" + "T" * 30 + "ACGTTGCA" * 30)
    print(f"[gpu-align] runs across chunks {preset} reads={len(reads)}", flush=True)
    got = ga.seeds(ctx, g, gi, reads, sd["density"], order=True)
    for s, b in zip(reads, got):
        a = oi.seeds(s, sd["density"], True, cap_per_base=512)
        assert a.shape == b.shape and (a == b).all()
    # the runs' k-mers are in this graph's index: thousands of seeds end in them
    assert sum(len(b) for b in got) > 3000


@pytest.mark.parametrize("mode", sorted(MODES))
def test_gpu_gaf_wide_fans(ctx, workdir, mode):
    """Slices across 48-way fans outgrow the band queue that a first attempt
    keeps in LDS: those extensions overflow and are rerun with the queue in
    their workspace, and the output still matches the oracle."""
    gfa, names, seqs = fan_workload(workdir)
    sd, og, oi, gp, g, gi = _both(gfa, "vg")
    _set_mode(gp, mode)
    ot, ostat, _ = o.align_reads(og, oi, o.preset_params("vg"), sd["density"], names, seqs, 8)
    print(f"[gpu-align] wide fans {mode} reads={len(seqs)}", flush=True)
    gt, gstat, st = ga.align_reads(ctx, g, gi, gp, names, seqs)
    assert list(gstat) == list(ostat)
    assert gt == ot
    if mode != "device_loop":
        assert st["reruns"] > 0


def test_gpu_gaf_buffer_matches_text(ctx, workdir):
    """The zero-copy GAF buffer holds the same bytes as the decoded text."""
    gfa, names, seqs = workload("mini", workdir)
    sd, og, oi, gp, g, gi = _both(gfa, "vg")
    print("[gpu-align] gaf buffer", flush=True)
    text, s1, _ = ga.align_reads(ctx, g, gi, gp, names, seqs)
    buf, s2, st = ga.align_reads(ctx, g, gi, gp, names, seqs, gaf_format="buffer")
    assert list(s1) == list(s2)
    assert buf.dtype.name == "uint8" and len(buf) == st["gafBytes"]
    assert buf.tobytes().decode() == text
    view = buf[:10]
    del buf
    assert view.tobytes().decode() == text[:10]
