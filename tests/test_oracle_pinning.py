"""Pins the oracle's extension (band DP, slice loop, traceback) against what the
mathematics fixes, beyond the reference's single known answer:

* with a bandwidth wider than the graph and X-drop off, the banded slice DP
  (src/GraphAlignerBitvectorBanded.h:254-725) computes every reachable column, so
  each slice's minimum score equals the minimum over all graph positions of the
  textbook edit-distance DP over the graph at the slice's last row (rows past
  the read end match nothing, as in getEqVector);
* a traceback (src/GraphAlignerBitvectorCommon.h:496-758) is a path through the
  graph whose edit count is at least the DP value it starts from (the score the
  extension returns: no path beats the optimum), and equals it unless the walk
  took one of the reference's out-of-band fallbacks (e.g. the horizontal walk to
  column 0 when the node was not in the previous slice, :630-637), which the
  test bounds to a small share of traces.

A wrong sign, index or dropped term in the column step, the merges, the band
relaxation or the backtrace choices breaks one of these.
"""
import collections
import os
import random

import numpy as np
import pytest

import oracle_ffi as o
from graphaligner_amd import synth
from workloads import revcomp

INF = 1 << 40


def _tiny_workload(workdir, name, genome, seed):
    prefix = os.path.join(workdir, name)
    if not os.path.exists(prefix + ".gfa"):
        synth.generate("mini", prefix, overrides_graph=dict(genomeLength=genome, seed=seed),
                       overrides_reads=dict(numReads=12, lenA=150, lenB=400, lenMin=150, lenMax=400, seed=seed + 1000))
    names, seqs = synth.read_fasta(prefix + ".fa")
    return prefix + ".gfa", seqs


class GraphView:
    def __init__(self, og):
        ex = og.export()
        self.len = ex["nodeLength"]
        self.ids = ex["nodeIDs"]
        self.off = ex["nodeOffset"]
        self.seq2 = ex["seq2"]
        self.outOff = ex["outOff"]
        self.outAdj = ex["outAdj"]
        self.firstAmbiguous = ex["firstAmbiguous"]
        self.pieces = collections.defaultdict(list)
        for n in range(len(self.len)):
            self.pieces[int(self.ids[n])].append(n)
        for k in self.pieces:
            self.pieces[k].sort(key=lambda n: int(self.off[n]))
        self.size = {k: sum(int(self.len[n]) for n in v) for k, v in self.pieces.items()}

    def char(self, n, p):
        assert n < self.firstAmbiguous
        w = int(self.seq2[2 * n + p // 32])
        return "ACGT"[(w >> (2 * (p % 32))) & 3]

    def out(self, n):
        return [int(x) for x in self.outAdj[self.outOff[n]:self.outOff[n + 1]]]

    def split_of(self, bigraph, offset):
        for n in self.pieces[bigraph]:
            if self.off[n] <= offset < self.off[n] + self.len[n]:
                return n, offset - int(self.off[n])
        raise ValueError


def brute_force_slice_minima(gv, start_node, start_off, seq):
    """Row-by-row edit-distance DP over the positions reachable from start_node;
    returns min over positions at rows 63, 127, ... (64-row slices)."""
    reach, stack = {start_node}, [start_node]
    while stack:
        n = stack.pop()
        for v in gv.out(n):
            if v not in reach:
                reach.add(v)
                stack.append(v)
    indeg = {n: 0 for n in reach}
    for n in reach:
        for v in gv.out(n):
            indeg[v] += 1
    order, queue = [], [n for n in reach if indeg[n] == 0]
    while queue:
        n = queue.pop()
        order.append(n)
        for v in gv.out(n):
            indeg[v] -= 1
            if indeg[v] == 0:
                queue.append(v)
    assert len(order) == len(reach), "the test graph must be acyclic"
    pos, index = [], {}
    for n in order:
        for p in range(int(gv.len[n])):
            index[(n, p)] = len(pos)
            pos.append((n, p))
    preds = [[] for _ in pos]
    for i, (n, p) in enumerate(pos):
        if p + 1 < gv.len[n]:
            preds[index[(n, p + 1)]].append(i)
        else:
            for v in gv.out(n):
                preds[index[(v, 0)]].append(i)
    chars = [gv.char(n, p) for n, p in pos]
    # row -1: the seed column scores 0, columns of the seed node score their distance to it
    prev = [INF] * len(pos)
    for i, (n, p) in enumerate(pos):
        if n == start_node:
            prev[i] = abs(p - start_off)
        for q in preds[i]:
            prev[i] = min(prev[i], prev[q] + 1)
    rows = 64 * ((len(seq) + 63) // 64)
    minima = []
    for r in range(rows):
        c = seq[r] if r < len(seq) else None
        cur = [INF] * len(pos)
        for i in range(len(pos)):
            best = prev[i] + 1
            mm = 0 if c == chars[i] else 1
            for q in preds[i]:
                best = min(best, cur[q] + 1, prev[q] + mm)
            cur[i] = best
        prev = cur
        if r % 64 == 63:
            minima.append(min(cur))
    return minima


def _cases(gv, seqs, n, seed, max_len):
    rng = random.Random(seed)
    keys = sorted(gv.pieces)
    out = []
    while len(out) < n:
        read = seqs[rng.randrange(len(seqs))]
        if rng.random() < 0.5:
            read = revcomp(read)
        L = rng.randrange(40, max_len)
        start = rng.randrange(0, max(1, len(read) - L))
        node = keys[rng.randrange(len(keys))]
        off = rng.randrange(gv.size[node])
        out.append((read[start:start + L], node, off))
    return out


def test_slice_minima_equal_brute_force_dp(workdir):
    gfa, seqs = _tiny_workload(workdir, "pin_tiny", 700, 31)
    og = o.Graph(gfa)
    gv = GraphView(og)
    p = o.preset_params("vg")
    p.bandwidth = 1 << 20
    p.xdrop = 1 << 30
    checked = 0
    for seq, node, off in _cases(gv, seqs, 6, 5, 150):
        rc, score, trace, rows = o.extend(og, p, seq, node, off)
        if rc != 0:
            continue
        sn, so = gv.split_of(node, off)
        want = brute_force_slice_minima(gv, sn, so, seq)
        got = [r[1] for r in rows]
        assert got == want[:len(got)]
        assert len(got) == len(want)
        checked += 1
    assert checked >= 4


def _recount(gv, seq, trace):
    """Edit count of a trace (rows: split node, offset, seqPos, nodeSwitch; end
    first), starting from the seed item, and checks that it is a graph path."""
    items = [tuple(int(x) for x in r) for r in trace[::-1]]
    assert items[0][2] == -1
    score = 0
    for (pn, po, ps, psw), (n, off, s, sw) in zip(items, items[1:]):
        assert s in (ps, ps + 1)
        if n == pn and off == po:
            assert s == ps + 1
            score += 1  # read base against no graph base
            continue
        if n == pn:
            assert off == po + 1
        else:
            assert po == gv.len[pn] - 1 and off == 0 and n in gv.out(pn)
        if s == ps:
            score += 1  # graph base against no read base
        elif seq[s] != gv.char(n, off):
            score += 1
    return score


@pytest.mark.parametrize("preset", ["mini", "mini_sv"])
def test_trace_edit_count_bounds_score(workdir, preset):
    from workloads import extension_cases, workload
    gfa, names, seqs = workload(preset, workdir)
    og = o.Graph(gfa)
    gv = GraphView(og)
    idx = o.Index(og, 15, 20, 0.001)
    p = o.preset_params("vg")
    n = exact = 0
    for seq, node, off in extension_cases(og, idx, seqs, 120, seed=3):
        rc, score, trace, rows = o.extend(og, p, seq, node, off)
        if rc != 0:
            continue
        r = _recount(gv, seq, trace)
        assert r >= score
        exact += r == score
        n += 1
    assert n >= 80
    assert exact >= 0.9 * n
