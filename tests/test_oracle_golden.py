"""Pins the oracle: the reference's own known answer and closed-form checks of the
bit-parallel column primitives (src/WordSlice.h, src/GraphAlignerBitvectorCommon.h)."""
import ctypes
import os
import random

import pytest

import oracle_ffi as o

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# README.md:30-40 runs `-x vg` on test/graph.gfa + test/read.fa. The README line was
# written by an older GAF writer ("M" cigar ops, no AS:f); the current writer
# (src/GraphAlignerGAFAlignment.h:156-199) splits M into =/X and adds
# AS:f = (71*100 - 294*6)/100 (src/GraphAligner.h:689-690). Same coordinates,
# path, NM, dv, id and op lengths as the README line.
README_LINE_OLD = "read\t71\t0\t71\t+\t>1>2>4\t87\t3\t73\t66\t72\t255\tNM:i:6\tdv:f:0.0833333\tid:f:0.916667\tcg:Z:4M1M2M1M1I37M1D5M1I5M1M13M"
README_LINE = "read\t71\t0\t71\t+\t>1>2>4\t87\t3\t73\t66\t72\t255\tNM:i:6\tAS:f:53.36\tdv:f:0.0833333\tid:f:0.916667\tcg:Z:4=1X2=1X1I37=1D5=1I5=1X13="


def read_fa(path):
    lines = open(path).read().split("\n")
    return lines[0][1:], lines[1].strip()


def test_readme_known_answer():
    g = o.Graph(os.path.join(GOLD, "readme_graph.gfa"))
    idx = o.Index(g, 15, 20, 0.001)
    name, seq = read_fa(os.path.join(GOLD, "readme_read.fa"))
    text, status, _ = o.align_reads(g, idx, o.preset_params("vg"), 10.0, [name], [seq])
    assert list(status) == [0]
    assert text == README_LINE + "\n"
    # the README's own line, modulo the =/X split and the AS:f field
    old = text.strip().replace("\tAS:f:53.36", "").split("\t")
    ref = README_LINE_OLD.split("\t")
    assert old[:-1] == ref[:-1]
    import re
    ops = re.findall(r"(\d+)([=XMID])", old[-1][5:])
    ref_ops = re.findall(r"(\d+)([=XMID])", ref[-1][5:])
    assert [(n, "M" if c in "=X" else c) for n, c in ops] == ref_ops


def _values(vp, vn, end):
    """row values -1..63 of a slice (WordSlice::getValue / getScoreBeforeStart)"""
    before = end - bin(vp).count("1") + bin(vn).count("1")
    vals = [before]
    for r in range(64):
        vals.append(vals[-1] + ((vp >> r) & 1) - ((vn >> r) & 1))
    return vals


def _rand_slice(rng, base=None):
    vp = vn = 0
    for r in range(64):
        x = rng.random()
        if x < 0.35:
            vp |= 1 << r
        elif x < 0.55:
            vn |= 1 << r
    if rng.random() < 0.1:
        vp, vn = (1 << 64) - 1, 0
    if rng.random() < 0.05:
        vp, vn = 0, (1 << 64) - 1
    before = rng.randrange(0, 200) if base is None else base + rng.randrange(-70, 70)
    end = before + bin(vp).count("1") - bin(vn).count("1")
    return vp, vn, end


def test_merge_is_cellwise_min():
    """mergeTwoSlices == cellwise minimum over rows -1..63, brute force."""
    L = o.lib()
    rng = random.Random(1)
    u64 = ctypes.c_uint64
    for _ in range(20000):
        a = _rand_slice(rng)
        b = _rand_slice(rng, base=_values(*a)[0])
        vp, vn, end = u64(), u64(), ctypes.c_int32()
        L.orc_merge(a[0], a[1], a[2], b[0], b[1], b[2], ctypes.byref(vp), ctypes.byref(vn), ctypes.byref(end))
        want = [min(x, y) for x, y in zip(_values(*a), _values(*b))]
        assert vp.value & vn.value == 0
        assert _values(vp.value, vn.value, end.value) == want


def test_difference_masks_exact():
    """differenceMasksBitTwiddle == exact per-row comparison (its debug twin,
    src/WordSlice.h:666-762, asserts the same for non-negative differences)."""
    L = o.lib()
    rng = random.Random(2)
    u64 = ctypes.c_uint64
    for _ in range(20000):
        a = _rand_slice(rng)
        b = _rand_slice(rng, base=_values(*a)[0])
        va, vb = _values(*a), _values(*b)
        ls, rs = u64(), u64()
        L.orc_difference_masks(a[0], a[1], b[0], b[1], vb[0] - va[0], ctypes.byref(ls), ctypes.byref(rs))
        want_l = sum(1 << r for r in range(64) if va[r + 1] < vb[r + 1])
        want_r = sum(1 << r for r in range(64) if va[r + 1] > vb[r + 1])
        assert (ls.value, rs.value) == (want_l, want_r)


def test_next_slice_is_dp_column():
    """getNextSlice == one column of the O(nm) edit-distance recurrence."""
    L = o.lib()
    rng = random.Random(3)
    u64 = ctypes.c_uint64
    for _ in range(5000):
        vp, vn, end = _rand_slice(rng)
        # valid DP columns have |delta| <= 1 everywhere; hin in {-1, 0, 1}
        prev = _values(vp, vn, end)
        eq = rng.getrandbits(64)
        hin = rng.choice([-1, 0, 1])
        new = [prev[0] + hin]
        for r in range(64):
            diag = prev[r] + (0 if (eq >> r) & 1 else 1)
            new.append(min(prev[r + 1] + 1, new[-1] + 1, diag))
        ovp, ovn, oend, ph, mh = u64(), u64(), ctypes.c_int32(), u64(), u64()
        L.orc_next_slice(eq, vp, vn, end, 1 if hin == 1 else 0, 1 if hin == -1 else 0, ctypes.byref(ovp), ctypes.byref(ovn),
                         ctypes.byref(oend), ctypes.byref(ph), ctypes.byref(mh))
        # the recurrence only applies to valid inputs: skip columns it cannot produce
        if any(abs(new[i + 1] - new[i]) > 1 for i in range(64)):
            continue
        assert _values(ovp.value, ovn.value, oend.value) == new
        hout = new[64] - prev[64]
        assert (ph.value >> 63) == (1 if hout == 1 else 0)
        assert (mh.value >> 63) == (1 if hout == -1 else 0)


def test_max_xscore_is_max_over_candidates():
    """maxXScoreLocalMinima == max over its candidate rows of (j+r+1)*100 - v(r)*cost;
    with extraMask all ones the candidates contain the argmax over all rows."""
    L = o.lib()
    rng = random.Random(4)
    for _ in range(5000):
        vp, vn, end = _rand_slice(rng)
        vals = _values(vp, vn, end)[1:]
        j = rng.randrange(0, 5000) * 64
        cost = rng.choice([294, 150, 1000])
        cells = rng.choice([64, rng.randrange(1, 65)])
        got = L.orc_max_xscore(vp, vn, end, j, cost, cells, (1 << 64) - 1)
        want = max((j + r + 1) * 100 - vals[r] * cost for r in range(cells))
        assert got == want
