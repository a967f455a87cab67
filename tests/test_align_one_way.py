"""AlignOneWay with the caller's seeds (src/GraphAlignerWrapper.h:55,
src/GraphAligner.h:94-181) and OrderSeeds (src/GraphAlignerWrapper.h:59): the
product's host pipeline (device code run on CPU threads) against the oracle,
result by result (scores, positions, X-scores, seed goodness, MAPQ, whole
traces), before and after selection, and OrderSeeds of seeds that carry no
alignment-graph node through the product's C ABI (host code, no GPU)."""
import numpy as np
import pytest

import oracle_ffi as O
import pipe_ffi as P
import workloads as W


def _setup(wl, preset, workdir):
    gfa, names, seqs = W.workload(wl, workdir)
    sd = O.PRESET_SEEDING[preset]
    og = O.Graph(gfa)
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    pg = P.Graph(gfa)
    pi = P.Index(pg, sd["k"], sd["w"], sd["discard"])
    return gfa, names, seqs, sd, og, oi, pi


@pytest.mark.parametrize("wl,preset,n", [("mini", "vg", 30), ("mini_sv", "vg", 30), ("mini_sv", "dbg", 30), ("vg100k", "vg", 20)])
def test_align_one_way_matches_oracle(workdir, wl, preset, n):
    gfa, names, seqs, sd, og, oi, pi = _setup(wl, preset, workdir)
    params = O.preset_params(preset)
    en, es = W.edge_reads(names, seqs)
    checked = 0
    for seq in list(seqs[:n]) + es:
        seeds = oi.seeds(seq, sd["density"], order=True)
        if seeds is None or len(seeds) == 0:
            continue
        for select, mapq in ((False, False), (True, True)):
            want = O.align_one_way(og, params, seq, seeds, select, mapq)
            got = P.align_one_way(pi, params, seq, seeds, select, mapq)
            assert got == want
        checked += 1
    assert checked >= n // 2


def test_caller_seeds_without_split_node(workdir):
    """Seeds from outside the minimizer seeder (no alignment-graph node, shuffled)
    ordered by the product's ga_order_seeds and by the oracle, then aligned."""
    import graphaligner_amd as ga
    gfa, names, seqs, sd, og, oi, pi = _setup("mini_sv", "vg", workdir)
    g = ga.Graph(gfa)
    params = O.preset_params("vg")
    rng = np.random.default_rng(5)
    for seq in seqs[:20]:
        seeds = oi.seeds(seq, sd["density"], order=False)
        if len(seeds) == 0:
            continue
        raw = seeds.copy()
        raw[:, 5:7] = -1
        raw[:, 8:10] = 0
        raw = raw[rng.permutation(len(raw))]
        want = O.order_seeds(og, raw)
        got = ga.order_seeds(g, raw)
        assert want is not None
        assert np.array_equal(got, want)
        # the same order as from the seeder's own split nodes
        assert np.array_equal(want[:, 8:10], O.order_seeds(og, raw)[:, 8:10])
        assert P.align_one_way(pi, params, seq, want, True, False) == O.align_one_way(og, params, seq, want, True, False)


def test_order_seeds_assertion_and_bad_node(workdir):
    import graphaligner_amd as ga
    gfa, names, seqs, sd, og, oi, pi = _setup("mini", "vg", workdir)
    g = ga.Graph(gfa)
    seeds = oi.seeds(seqs[0], sd["density"], order=False).copy()
    # a read position beyond the seed's chain position: the reference asserts
    # (src/GraphAligner.h:270,276)
    bad = seeds.copy()
    bad[0, 2] = 10 ** 9
    assert O.order_seeds(og, bad) is None
    with pytest.raises(RuntimeError, match="GA_ASSERT"):
        ga.order_seeds(g, bad)
    nonode = seeds.copy()
    nonode[0, 0] = 10 ** 8
    with pytest.raises(RuntimeError, match="GA_EBADARG"):
        ga.order_seeds(g, nonode)


def test_align_one_way_edge_cases(workdir):
    gfa, names, seqs, sd, og, oi, pi = _setup("mini", "vg", workdir)
    params = O.preset_params("vg")
    seq = seqs[3]
    seeds = oi.seeds(seq, sd["density"], order=True)
    # no seeds: AlignOneWay asserts (src/GraphAligner.h:99)
    empty = np.zeros((0, 10), dtype=np.int64)
    assert O.align_one_way(og, params, seq, empty) == P.align_one_way(pi, params, seq, empty) == "S 2 0\n"
    # a first seed past the read's end: the assertion of getTwoDirectionalTrace
    past = seeds.copy()
    past[0, 2] = len(seq) + 5
    assert O.align_one_way(og, params, seq, past).startswith("S 2")
    assert P.align_one_way(pi, params, seq, past) == O.align_one_way(og, params, seq, past)
    # the same seed last, after a seed whose alignment covers the read end to end:
    # the sloppy loop stops before it (no assertion)
    tail = np.concatenate([seeds, past[:1]])
    tail[-1, 8] = 0
    assert P.align_one_way(pi, params, seq, tail) == O.align_one_way(og, params, seq, tail)
    # speculation on: a seed past the end picked ahead must not fail the read
    for spec in (1, 64):
        assert P.align_one_way(pi, params, seq, tail, speculative=spec) == O.align_one_way(og, params, seq, tail)
