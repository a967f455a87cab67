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


def edge_reads(names, seqs, seed=11):
    """Reads exercising the per-read outcomes of src/Aligner.cpp:390-606: no seeds
    (empty, shorter than k, random), reverse strand, lowercase, N runs, IUPAC
    letters, a letter the complement table does not know (assertion -> dropped),
    and a chimera of two distant reads (several alignments per read)."""
    rng = random.Random(seed)
    base = [s for s in seqs if len(s) > 2000][:4]
    out = []

    def add(name, s):
        out.append((name, s))

    add("empty", "")
    add("short", "ACGTACGTAC")
    add("random", "".join(rng.choice("ACGT") for _ in range(3000)))
    add("revcomp", revcomp(base[0]))
    add("lower", base[1].lower())
    s = base[2]
    add("nrun", s[:700] + "N" * 150 + s[850:])
    add("iupac", s[:500] + "R" + s[501:1200] + "y" + s[1201:])
    add("badletter", s[:900] + "X" + s[901:])
    add("chimera", base[0][:1500] + revcomp(base[3][-1500:]))
    add("tail_only", base[1][-200:])
    return [n for n, _ in out], [s for _, s in out]


def fan_workload(workdir, fan=48, regions=12, seed=5):
    """A chain of random segments joined by fans of `fan` parallel one-base nodes:
    a slice that crosses a fan holds more band nodes and queued edges than the
    first-attempt capacities of an extension's band queue, so those extensions
    overflow and are rerun with larger workspaces."""
    path = os.path.join(workdir, f"fan{fan}.gfa")
    rng = random.Random(seed)
    segs = ["".join(rng.choice("ACGT") for _ in range(rng.randrange(150, 400))) for _ in range(regions + 1)]
    lines = ["H\tVN:Z:1.0"]
    nid = 0
    seg_id = []
    fan_ids = []
    for r, s in enumerate(segs):
        nid += 1
        seg_id.append(nid)
        lines.append(f"S\t{nid}\t{s}")
        if r < regions:
            ids = []
            for k in range(fan):
                nid += 1
                ids.append(nid)
                lines.append(f"S\t{nid}\t{'ACGT'[k % 4]}")
            fan_ids.append(ids)
    for r in range(regions):
        for f in fan_ids[r]:
            lines.append(f"L\t{seg_id[r]}\t+\t{f}\t+\t0M")
            lines.append(f"L\t{f}\t+\t{seg_id[r + 1]}\t+\t0M")
    with open(path, "w") as out:
        out.write("\n".join(lines) + "\n")
    genome = "".join(s + (rng.choice("ACGT") if r < regions else "") for r, s in enumerate(segs))
    names, seqs = [], []
    for i in range(40):
        a = rng.randrange(0, len(genome) - 500)
        s = list(genome[a:a + rng.randrange(400, min(2000, len(genome) - a))])
        for k in range(len(s)):
            if rng.random() < 0.05:
                s[k] = rng.choice("ACGT")
        seq = "".join(s)
        names.append(f"fan{i}")
        seqs.append(seq if i % 2 == 0 else revcomp(seq))
    return path, names, seqs


def selfloop_workload(workdir, seed=9):
    """Backbone segments joined through tandem-repeat nodes that carry an edge to
    themselves (L u + u + 0M): a read crossing a tandem array loops through the
    same node several times within one 64-row slice, so its traceback re-enters
    a node of the current slice at a column right of where it entered it."""
    path = os.path.join(workdir, f"selfloop{seed}.gfa")
    rng = random.Random(seed)
    lines = ["H\tVN:Z:1.0"]
    units, segs = [], []
    nid = 0
    prev = None
    for r in range(8):
        nid += 1
        seg = "".join(rng.choice("ACGT") for _ in range(rng.randrange(300, 700)))
        lines.append(f"S\t{nid}\t{seg}")
        segs.append(seg)
        if prev is not None:
            lines.append(f"L\t{prev}\t+\t{nid}\t+\t0M")
        prev = nid
        if r < 7:
            nid += 1
            unit = "".join(rng.choice("ACGT") for _ in range(rng.randrange(12, 45)))
            lines.append(f"S\t{nid}\t{unit}")
            lines.append(f"L\t{prev}\t+\t{nid}\t+\t0M")
            lines.append(f"L\t{nid}\t+\t{nid}\t+\t0M")
            units.append(unit)
            prev = nid
    with open(path, "w") as out:
        out.write("\n".join(lines) + "\n")
    names, seqs = [], []
    for i in range(30):
        parts = []
        for r in range(8):
            parts.append(segs[r])
            if r < 7:
                parts.append(units[r] * rng.randrange(2, 12))
        genome = "".join(parts)
        a = rng.randrange(0, len(genome) // 2)
        s = list(genome[a:a + rng.randrange(800, 2500)])
        for k in range(len(s)):
            if rng.random() < 0.06:
                s[k] = rng.choice("ACGT")
        seq = "".join(s)
        names.append(f"loop{i}")
        seqs.append(seq if i % 2 == 0 else revcomp(seq))
    return path, names, seqs
