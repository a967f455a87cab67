"""The product's host-built minimizer index, exported through ga_index_export,
against the oracle's index (MinimizerSeeder, src/MinimizerSeeder.cpp:299-491):
the same k-mers, each with the same positions in the same storage order, and
ranges that tile the position array. The GPU build is compared with this one
in test_gpu_index.py."""
import os

import numpy as np
import pytest

import oracle_ffi as O
import graphaligner_amd as ga
import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden")
IUPAC_GFA = ("S\t1\tACGTACGTACGGATTACAGATTACA\nS\t2\tACNTRYACGATTACCAGGATTACAAC\nS\t3\tGGGTTTCCCAGATTACAGGCATTACA\n"
             "L\t1\t+\t2\t+\t0M\nL\t2\t+\t3\t+\t0M\nL\t1\t+\t3\t+\t0M\n")


def graphs(workdir):
    out = []
    for wl, preset in (("mini_sv", "vg"), ("mini_dbg", "dbg")):
        gfa, _, _ = W.workload(wl, workdir)
        out.append((gfa, preset, None))
    path = os.path.join(workdir, "iupac_index.gfa")
    with open(path, "w") as f:
        f.write(IUPAC_GFA)
    out.append((path, "vg", (5, 9)))
    out.append((os.path.join(GOLD, "readme_graph.gfa"), "vg", (7, 11)))
    return out


def product_params(preset, kw):
    p = ga.params_preset(preset)
    if kw:
        p.minimizerLength, p.minimizerWindowSize = kw
    return p


def exported_table(index):
    kmers, starts, counts, pos = index.export()
    assert np.all(np.diff(kmers.astype(np.float64)) > 0) if len(kmers) > 1 else True
    order = np.argsort(starts, kind="stable")
    # the ranges tile [0, len(pos)) in increasing start order
    assert int(counts.sum()) == len(pos)
    assert np.array_equal(starts[order], np.concatenate([[0], np.cumsum(counts[order])[:-1]]).astype(np.int64)) if len(order) else True
    return {int(k): pos[s: s + c].tolist() for k, s, c in zip(kmers, starts, counts)}


@pytest.mark.parametrize("case", range(4))
def test_host_index_matches_oracle(workdir, case):
    gfa, preset, kw = graphs(workdir)[case]
    p = product_params(preset, kw)
    k, w = p.minimizerLength, p.minimizerWindowSize
    og = O.Graph(gfa)
    oi = O.Index(og, k, w, p.minimizerDiscardMostNumerousFraction)
    gi = ga.Index(ga.Graph(gfa), p)
    want = O.index_table(oi)
    got = exported_table(gi)
    assert got == want
    assert gi.num_kmers == len(want)
    assert gi.max_count == oi.max_count
