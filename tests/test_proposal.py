"""The learned proposal's host side on the CPU (the product's host pipeline
over device code run on host threads): seed clusters agree with OrderSeeds
(src/GraphAligner.h:250-312), cluster features are finite and bounded,
pruning removes exactly the clusters the scorer rates below the threshold
(except the first keepTop), a threshold nothing falls below leaves the GAF
byte for byte that of the plain method, and the outcome records label every
cluster by the alignments its seeds made."""
import numpy as np
import pytest

import oracle_ffi as O
import pipe_ffi as P
import workloads as W


@pytest.fixture(scope="module")
def setup(workdir):
    gfa, names, seqs = W.workload("mini_sv", workdir)
    sd = O.PRESET_SEEDING["vg"]
    pg = P.Graph(gfa)
    pi = P.Index(pg, sd["k"], sd["w"], sd["discard"])
    return names, seqs, sd, pi


def test_clusters_agree_with_order_seeds(setup):
    names, seqs, sd, pi = setup
    checked = 0
    for seq in seqs[:30]:
        seeds = pi.seeds(seq, sd["density"], order=True)
        if seeds is None or len(seeds) == 0:
            continue
        cl, feats = P.cluster_features(pi, seq, seeds)
        sizes = np.bincount(cl)
        # OrderSeeds' cluster size of every seed, and the covered bases it adds
        # to every seed of a cluster (goodness - raw goodness), agree
        assert np.array_equal(sizes[cl], seeds[:, 9])
        cover = seeds[:, 8] - seeds[:, 7]
        for c in range(len(sizes)):
            assert len(set(cover[cl == c])) == 1
        # numbered by first appearance in the seed order
        first = [int(np.argmax(cl == c)) for c in range(len(sizes))]
        assert first == sorted(first)
        assert np.isfinite(feats).all() and feats.min() >= 0 and feats.max() <= 4
        assert (feats[:, 47] == 1).all() and feats[0, 37] == 1 and (feats[1:, 37] == 0).all()
        checked += 1
    assert checked >= 20


def _plain(pi, names, seqs, sd):
    P.set_linear_proposal(None, 0, 0)
    return P.align_reads(pi, O.preset_params("vg"), sd["density"], names, seqs)


def test_threshold_below_every_score_is_plain(setup):
    names, seqs, sd, pi = setup
    plain, pst, _ = _plain(pi, names, seqs, sd)
    try:
        P.set_linear_proposal(np.zeros(48), 0.0, 0.5)  # every cluster scores 0.5
        text, st, _ = P.align_reads(pi, O.preset_params("vg"), sd["density"], names, seqs)
        s = P.proposal_stats()
    finally:
        P.set_linear_proposal(None, 0, 0)
    assert text == plain and np.array_equal(st, pst)
    assert s["clusters"] > len(seqs) and s["pruned_clusters"] == 0 and s["pruned_seeds"] == 0


@pytest.mark.parametrize("keep_top", [1, 2])
def test_pruning_counts_match_the_scores(setup, keep_top):
    """The scorer's rule applied in numpy to each read's clusters predicts the
    clusters and seeds removed; reads with nothing removed keep their GAF."""
    names, seqs, sd, pi = setup
    rng = np.random.default_rng(4)
    w = rng.normal(0, 1.0, 48).astype(np.float32)
    bias, thr = -0.5, 0.45
    want_clusters = want_pruned = want_seeds = 0
    untouched = []
    for i, seq in enumerate(seqs):
        seeds = pi.seeds(seq, sd["density"], order=True)
        if seeds is None or len(seeds) == 0:
            continue
        cl, feats = P.cluster_features(pi, seq, seeds)
        z = feats.astype(np.float64) @ w.astype(np.float64) + bias
        p = (1.0 / (1.0 + np.exp(-z))).astype(np.float32)
        assert np.abs(p.astype(np.float64) - thr).min() > 1e-5  # no score on the threshold
        drop = (p < thr) & (np.arange(len(p)) >= keep_top)
        want_clusters += len(p)
        want_pruned += int(drop.sum())
        want_seeds += int(drop[cl].sum())
        if not drop.any():
            untouched.append(i)
    plain, pst, _ = _plain(pi, names, seqs, sd)
    try:
        P.set_linear_proposal(w, bias, thr, keep_top)
        text, st, _ = P.align_reads(pi, O.preset_params("vg"), sd["density"], names, seqs)
        s = P.proposal_stats()
    finally:
        P.set_linear_proposal(None, 0, 0)
    assert s == dict(clusters=want_clusters, pruned_clusters=want_pruned, pruned_seeds=want_seeds)
    assert want_pruned > 0 and untouched

    def lines_of(t, name):
        return [ln for ln in t.split("\n") if ln.split("\t", 1)[0] == name]
    for i in untouched:
        assert lines_of(text, names[i]) == lines_of(plain, names[i])
        assert st[i] == pst[i]


def test_records_label_the_clusters(setup):
    names, seqs, sd, pi = setup
    P.records(True)
    try:
        text, st, _ = P.align_reads(pi, O.preset_params("vg"), sd["density"], names, seqs)
        F, Y, R = P.take_records()
    finally:
        P.records(False)
    plain, pst, _ = _plain(pi, names, seqs, sd)
    assert text == plain  # recording changes nothing
    expect = 0
    for i, seq in enumerate(seqs):
        seeds = pi.seeds(seq, sd["density"], order=True)
        if seeds is not None and len(seeds):
            cl, feats = P.cluster_features(pi, seq, seeds)
            rows = R == i
            if st[i] in (0, 3):
                assert rows.sum() == len(feats)
                assert np.array_equal(F[rows], feats)
                expect += len(feats)
    assert len(R) == expect and np.all(np.diff(R) >= 0)
    assert (Y[:, 0] <= Y[:, 1]).all()  # kept implies made
    for i in np.unique(R):
        assert Y[R == i, 0].any() == (st[i] == 0)


def test_clusters_past_the_scored_ones_are_removed(setup):
    """Only a read's first max_clusters clusters are scored; the rest go unscored."""
    names, seqs, sd, pi = setup
    want_scored = want_pruned = 0
    for seq in seqs:
        seeds = pi.seeds(seq, sd["density"], order=True)
        if seeds is None or len(seeds) == 0:
            continue
        cl, _ = P.cluster_features(pi, seq, seeds)
        c = int(cl.max()) + 1
        want_scored += min(c, 1)
        want_pruned += max(0, c - 1)
    try:
        P.set_linear_proposal(np.zeros(48), 0.0, 0.5, 1, max_clusters=1)  # every scored cluster passes
        P.align_reads(pi, O.preset_params("vg"), sd["density"], names, seqs)
        s = P.proposal_stats()
    finally:
        P.set_linear_proposal(None, 0, 0)
    assert s["clusters"] == want_scored and s["pruned_clusters"] == want_pruned and want_pruned > 0
