"""Pins the oracle's slice DP at a finite band, its minimizer index, OrderSeeds and
SelectAlignments against checks that do not restate their code:

* the reference's own debug-only invariant checks (compiled in the reference
  under EXTRACORRECTNESSASSERTIONS): assertSliceCorrectness on every column step
  (src/GraphAlignerBitvectorCommon.h:1062-1076, called at :1447-1448) and
  checkNodeBoundaryCorrectness on every slice (src/GraphAlignerBitvectorBanded.h:
  169-227, called at :492-494), restated in an oracle build that counts
  violations; run at b = 10 and b = 5, with and without a tangle-effort cap, on
  variation graphs with inversions and on the cyclic de Bruijn graph;
* the minimizers of iterateMinimizersReal against a statement of them through
  window minima (and their relation to the reference's brute-force twin
  iterateMinimizersSimple, src/MinimizerSeeder.cpp:201-265), and the whole
  index (positions per k-mer and maxCount, :299-491, :557-575) against a
  count table built from that statement;
* OrderSeeds (src/GraphAligner.h:250-312) against a set-union formulation of
  its cluster coverage;
* SelectAlignments (src/AlignmentSelection.cpp:13-95) against an O(n^2) filter.

And measures how much the output depends on the band queue's tie order inside
strongly connected components (the order phmap would give is unavailable): the
oracle built with ties broken the other way, GAF lines compared.
"""
import random

import numpy as np
import pytest

import oracle_ffi as O
import workloads as W

M64 = (1 << 64) - 1


def _hash(key):
    """the integer mix of src/MinimizerSeeder.cpp:45-54, on 64-bit words"""
    key = (~key + (key << 21)) & M64
    key ^= key >> 24
    key = (key + (key << 3) + (key << 8)) & M64
    key ^= key >> 14
    key = (key + (key << 2) + (key << 4)) & M64
    key ^= key >> 28
    key = (key + (key << 31)) & M64
    return key


CODE = {"A": 0, "C": 1, "G": 2, "T": 3, "a": 0, "c": 1, "g": 2, "t": 3}


def _minimizers_simple(s, k, w):
    """iterateMinimizersSimple (src/MinimizerSeeder.cpp:201-265): every k-mer whose
    hash is the minimum of some window of w bases (of the whole string when it is
    at most w long); (position of the k-mer's last base, k-mer), by position."""
    n = len(s)
    if n < k:
        return []
    kmer = [0] * n
    h = [None] * n
    for i in range(k - 1, n):
        v = 0
        for c in s[i + 1 - k: i + 1]:
            v = (v << 2) | CODE[c]
        kmer[i] = v
        h[i] = _hash(v)
    if n <= w:
        m = min(h[k - 1:])
        return [(i, kmer[i]) for i in range(k - 1, n) if h[i] == m]
    mark = [False] * n
    for i in range(w - 1, n):
        lo = i + k - w
        m = min(h[lo: i + 1])
        for j in range(lo, i + 1):
            if h[j] == m:
                mark[j] = True
    return [(i, kmer[i]) for i in range(n) if mark[i]]


def _kmers(s, k):
    km, h = {}, {}
    for i in range(k - 1, len(s)):
        v = 0
        for c in s[i + 1 - k: i + 1]:
            v = (v << 2) | CODE[c]
        km[i] = v
        h[i] = _hash(v)
    return km, h


def _minimizers_by_windows(s, k, w):
    """What iterateMinimizersReal (src/MinimizerSeeder.cpp:104-189) reports,
    stated through window minima rather than its deque. With r = w - k + 1
    k-mers per window: the first window holds r + 1 k-mers (ending at k-1 ..
    k-1+r) and all its minimum positions are reported; each later window ends at
    e and holds the r k-mers ending at e-r+1 .. e. When its minimum m is below
    the previous window's, e is reported; above it (the old minimum left), the
    last earlier position holding m and e if it holds m; equal, e if it holds m.
    A string of at most w bases reports nothing (the first window reads one
    base past it)."""
    n = len(s)
    r = w - k + 1
    if n <= w:
        return []
    km, h = _kmers(s, k)
    out = set()
    first = range(k - 1, k + r)
    m = min(h[j] for j in first)
    out.update(j for j in first if h[j] == m)
    prev = m
    for e in range(k + r, n):
        win = range(e - r + 1, e + 1)
        m = min(h[j] for j in win)
        if m < prev:
            out.add(e)
        elif m > prev:
            older = [j for j in win if j < e and h[j] == m]
            if older:
                out.add(max(older))
            if h[e] == m:
                out.add(e)
        elif h[e] == m:
            out.add(e)
        prev = m
    return sorted((j, km[j]) for j in out)


def _revcomp(s):
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


def _gfa_nodes(gfa):
    seqs = {}
    for line in open(gfa):
        if line.startswith("S\t"):
            f = line.rstrip("\n").split("\t")
            seqs[f[1]] = f[2]
    return seqs


# ---------------------------------------------------------------------------


@pytest.mark.parametrize("wl,preset,cap", [("mini_sv", "vg", 0), ("mini_sv", "dbg", 0), ("mini_sv", "vg", 200), ("mini_dbg", "dbg", 0),
                                           ("mini_dbg", "dbg", 400), ("vg100k", "vg", 0)])
def test_reference_invariants_at_finite_band(workdir, wl, preset, cap):
    gfa, names, seqs = W.workload(wl, workdir)
    og = O.Graph(gfa, ndebug="checked")
    sd = O.PRESET_SEEDING[preset]
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    p = O.preset_params(preset)
    if cap:
        p.maxCellsPerSlice = cap
    L = og.L
    run0, fail0 = L.orc_extra_checks_run(), L.orc_extra_check_failures()
    notvalid = 0
    for c in W.extension_cases(og, oi, seqs, 50, seed=3, density=sd["density"]):
        r = O.extend(og, p, *c)
        notvalid += sum(x[9] for x in r[3])
    runs, fails = L.orc_extra_checks_run() - run0, L.orc_extra_check_failures() - fail0
    assert runs > 1_000_000
    assert fails == 0
    if cap:
        assert notvalid > 0


@pytest.mark.parametrize("k,w", [(15, 20), (19, 30), (5, 9), (3, 3)])
def test_minimizers_by_window_minima(workdir, k, w):
    """The oracle's minimizers equal the window-minimum statement above, and
    relate to the reference's brute-force twin iterateMinimizersSimple as its
    (commented-out) comparison at src/MinimizerSeeder.cpp:276-281 leaves them:
    every reported position is also the twin's, apart from the first window's
    extra k-mer (the twin's windows all hold w - k + 1 k-mers)."""
    rng = random.Random(k * 100 + w)
    strings = ["".join(rng.choice("ACGT") for _ in range(rng.randrange(1, 400))) for _ in range(150)]
    # low-complexity strings repeat k-mers, so window minima tie
    strings += ["".join(rng.choice("AC") for _ in range(rng.randrange(w, 300))) for _ in range(60)]
    strings += ["A" * n for n in (k - 1, k, w - 1, w, w + 1, 3 * w)]
    gfa, _, _ = W.workload("mini_sv", workdir)
    for s in list(_gfa_nodes(gfa).values())[:300]:
        strings += [s, _revcomp(s)]
    extra_first = 0
    for s in strings:
        got = sorted(map(tuple, O.minimizers(s, k, w).tolist()))
        assert got == _minimizers_by_windows(s, k, w), (s, k, w)
        outside = set(got) - set(_minimizers_simple(s, k, w))
        assert all(p == k - 1 + (w - k + 1) for p, _ in outside)
        extra_first += len(outside)


@pytest.mark.parametrize("wl,preset", [("mini_sv", "vg"), ("mini_dbg", "dbg")])
def test_index_equals_count_table(workdir, wl, preset):
    gfa, _, _ = W.workload(wl, workdir)
    sd = O.PRESET_SEEDING[preset]
    og = O.Graph(gfa)
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    ex = og.export()
    node_ids, node_off, node_len = ex["nodeIDs"].tolist(), ex["nodeOffset"].tolist(), ex["nodeLength"].tolist()
    in_off, in_adj = ex["inOff"].tolist(), ex["inAdj"].tolist()
    # split nodes of each bigraph node by offset, and where its minimizers start:
    # at the last split node whose in-neighbours include another bigraph node
    parts, start = {}, {}
    for i, nid in enumerate(node_ids):
        parts.setdefault(nid, []).append((node_off[i], i))
        start.setdefault(nid, 0)
        if any(node_ids[n] != nid for n in in_adj[in_off[i]: in_off[i + 1]]):
            start[nid] = max(start[nid], node_off[i])
    want = {}
    for nid, plist in parts.items():
        plist.sort()
        # the bigraph node's sequence from its split nodes' 2-bit chunks
        seq = []
        for off, i in plist:
            w0, w1 = int(ex["seq2"][2 * i]), int(ex["seq2"][2 * i + 1])
            for o in range(node_len[i]):
                seq.append("ACGT"[((w0 if o < 32 else w1) >> ((o & 31) * 2)) & 3])
        seq = "".join(seq)
        for pos, km in _minimizers_by_windows(seq, sd["k"], sd["w"]):
            if pos < start[nid]:
                continue
            off, split = max((o, i) for o, i in plist if o <= pos)
            want.setdefault(km, []).append((split << 6) | (pos - off))
    assert ex["firstAmbiguous"] == len(node_ids)  # no IUPAC nodes: 2-bit chunks hold every base
    got = O.index_table(oi)
    assert set(got) == set(want)
    for km in want:
        assert sorted(got[km]) == sorted(want[km]), km
    # maxCount: one past the count at the kept fraction of the sorted counts, the
    # smallest count left out (the BBHash key initMaxCount skips, substituted)
    counts = sorted(len(v) for v in want.values())[1:]
    idx = min(int(len(counts) * (1.0 - sd["discard"])), len(counts) - 1)
    assert oi.max_count == counts[idx] + 1


def _order_brute(g, seeds):
    """OrderSeeds' goodness and cluster sizes, computed as sets: per chain, seeds
    sorted by diagonal split where consecutive diagonals differ by more than 100;
    a cluster's covered bases are the union of (seqPos - matchLen + 1, seqPos]."""
    ex = g
    res = {}
    by_chain = {}
    for idx, s in enumerate(seeds):
        node, off, pos = int(s[5]), int(s[6]), int(s[2])
        by_chain.setdefault(int(ex["chain"][node]), []).append((int(ex["chainPos"][node]) + off - pos, idx))
    for items in by_chain.values():
        items.sort()
        clusters, cur = [], [items[0]]
        for a, b in zip(items, items[1:]):
            if b[0] <= a[0] + 100:
                cur.append(b)
            else:
                clusters.append(cur)
                cur = [b]
        clusters.append(cur)
        for cl in clusters:
            covered = set()
            for _, idx in cl:
                pos, ml = int(seeds[idx][2]), int(seeds[idx][3])
                covered.update(range(pos - ml + 2, pos + 1))
            for _, idx in cl:
                res[idx] = (len(covered) + int(seeds[idx][7]), len(cl))
    return res


@pytest.mark.parametrize("wl,preset", [("mini_sv", "vg"), ("mini_dbg", "dbg"), ("vg100k", "vg")])
def test_order_seeds_against_set_formulation(workdir, wl, preset):
    gfa, names, seqs = W.workload(wl, workdir)
    sd = O.PRESET_SEEDING[preset]
    og = O.Graph(gfa)
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    ex = og.export()
    checked = 0
    for s in seqs[:25]:
        raw = oi.seeds(s, sd["density"], order=False)
        if len(raw) == 0:
            continue
        ordered = O.order_seeds(og, raw)
        assert ordered is not None
        want = _order_brute(ex, raw)
        # the same seeds, each with the brute-force goodness and cluster size
        key = lambda r: tuple(int(x) for x in r[:8])
        exp = sorted(key(raw[i]) + want[i] for i in range(len(raw)))
        assert sorted(key(r) + (int(r[8]), int(r[9])) for r in ordered) == exp
        # best first
        assert all(a >= b for a, b in zip(ordered[:, 8], ordered[1:, 8]))
        checked += 1
    assert checked > 10


def _parse_one_way(text):
    lines = text.split("\n")
    out = []
    for i, ln in enumerate(lines):
        if ln.startswith("A "):
            f = ln.split()
            out.append((int(f[1]), int(f[2]), float(f[4]), lines[i + 1]))
    return out


def _select_brute(alns, min_score, fraction):
    kept = [a for a in alns if a[2] >= int(min_score)]
    f32 = np.float32
    result = []
    for i, a in enumerate(kept):
        drop = False
        for j, b in enumerate(kept):
            if i == j:
                continue
            shorter = min(a[1] - a[0], b[1] - b[0])
            lo, hi = (a, b) if a[0] <= b[0] else (b, a)
            overlap = max(0, lo[1] - hi[0])
            if f32(overlap) > f32(shorter) * f32(0.05) and a[2] < b[2] * fraction:
                drop = True
                break
        if not drop:
            result.append(a)
    return result


@pytest.mark.parametrize("wl,preset", [("mini_sv", "vg"), ("mini_dbg", "dbg")])
def test_select_alignments_against_quadratic_filter(workdir, wl, preset):
    gfa, names, seqs = W.workload(wl, workdir)
    en, es = W.edge_reads(names, seqs)
    sd = O.PRESET_SEEDING[preset]
    og = O.Graph(gfa)
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    p = O.preset_params(preset)
    p.sloppy = 0  # try every seed: more, overlapping alignments per read
    multi = 0
    for s in list(seqs[:20]) + es:
        seeds = oi.seeds(s, sd["density"], order=True)
        if seeds is None or len(seeds) == 0:
            continue
        raw = _parse_one_way(O.align_one_way(og, p, s, seeds))
        sel = _parse_one_way(O.align_one_way(og, p, s, seeds, select=True))
        want = _select_brute(raw, 0, p.multimapScoreFraction)
        assert sorted(sel) == sorted(want)
        assert all(a[2] >= b[2] for a, b in zip(sel, sel[1:]))
        multi += len(raw) > len(sel)
    assert multi > 0


def test_tie_order_sensitivity(workdir, capsys):
    """GAF lines that change when the band queue's ties inside strongly connected
    components are broken the other way (printed; the test records them)."""
    report = []
    for wl, preset, n in [("mini_sv", "vg", 40), ("vg100k", "vg", 200), ("mini_dbg", "dbg", 60)]:
        gfa, names, seqs = W.workload(wl, workdir)
        names, seqs = names[:n], seqs[:n]
        sd = O.PRESET_SEEDING[preset]
        out = []
        for variant in (False, "revties"):
            og = O.Graph(gfa, ndebug=variant)
            oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
            out.append(O.align_reads(og, oi, O.preset_params(preset), sd["density"], names, seqs, 8))
        a = out[0][0].split("\n")
        b = set(out[1][0].split("\n"))
        changed = sum(1 for ln in a if ln and ln not in b)
        report.append((wl, preset, len([x for x in a if x]), changed, int((out[0][1] != out[1][1]).sum())))
    with capsys.disabled():
        for r in report:
            print("[tie order] %s -x %s: %d GAF lines, %d change, %d read statuses change" % r)
    # acyclic graphs have no ties inside components: nothing may change there
    assert report[1][3] == 0
