"""Shared test workloads: synthetic graphs/reads and random extension cases."""
import collections
import os
import random

from graphaligner_amd import synth

_cache = {}


def revcomp(s):
    return s[::-1].translate(str.maketrans("ACGTacgt", "TGCAtgca"))


def workload(preset, workdir):
    key = (preset, workdir)
    if key not in _cache:
        prefix = os.path.join(workdir, preset)
        synth.generate(preset, prefix)
        names, seqs = synth.read_fasta(prefix + ".fa")
        _cache[key] = (prefix + ".gfa", names, seqs)
    return _cache[key]


def extension_cases(oracle_graph, oracle_index, seqs, n, seed=7, density=10.0):
    """(sequence, bigraph node id, offset) for both directions of random seeds,
    built the way GraphAligner::getTwoDirectionalTrace (src/GraphAligner.h:475-520) does."""
    ex = oracle_graph.export()
    osize = collections.Counter()
    for nid, ln in zip(ex["nodeIDs"], ex["nodeLength"]):
        osize[int(nid)] += int(ln)
    rng = random.Random(seed)
    cases = []
    tries = 0
    while len(cases) < n and tries < 20 * n:
        tries += 1
        seq = seqs[rng.randrange(len(seqs))]
        seeds = oracle_index.seeds(seq, density)
        if len(seeds) == 0:
            continue
        s = seeds[rng.randrange(len(seeds))]
        node_id, node_offset, seq_pos, rev = int(s[0]), int(s[1]), int(s[2]), int(s[4])
        fwd = node_id * 2 + rev
        if seq_pos < len(seq) - 1:
            cases.append((seq[seq_pos + 1:], fwd, node_offset))
        if seq_pos > 0:
            rc = revcomp(seq)
            cases.append((rc[len(rc) - seq_pos:], node_id * 2 + (1 - rev), osize[fwd] - node_offset - 1))
    return cases[:n]
