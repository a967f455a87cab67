"""ga_extend_batch_noalloc (the entry point that allocates no device memory:
its buffers are carved from a file-scope device array) against ga_extend_batch
through the C ABI: the same statuses, scores and traces on two calls in a row,
the second smaller, so the shared scratch's leftovers from the first call
cannot leak into the second; plus the size limit's error."""
import numpy as np
import pytest

import graphaligner_amd as ga
import oracle_ffi as o
from workloads import extension_cases, workload

pytestmark = pytest.mark.gpu


def _check_same(a, b):
    assert len(a) == len(b)
    for (sa, ca, ta, _), (sb, cb, tb) in zip(a, b):
        assert sa == sb and ca == cb
        assert np.array_equal(ta, tb)


def test_noalloc_matches_extend_batch(workdir):
    gfa, names, seqs = workload("vg100k", workdir)
    og = o.Graph(gfa)
    idx = o.Index(og, 15, 20, 0.001)
    cases = extension_cases(og, idx, seqs, 160, seed=3)
    ctx = ga.Context(0)
    g = ga.Graph(gfa)
    p = ga.params_preset("vg")
    # the graph must be resident first
    with pytest.raises(RuntimeError, match="GA_EBADARG"):
        ga.extend_batch_noalloc(ctx, g, p, [cases[0][0]], [cases[0][1]], [cases[0][2]])
    print("[no-alloc] reference call", flush=True)
    want, _ = ga.extend_batch(ctx, g, p, [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases])
    assert sum(1 for w in want if w[0] == 0) > 100
    print(f"[no-alloc] call 1: {len(cases)} extensions", flush=True)
    got, ms = ga.extend_batch_noalloc(ctx, g, p, [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases])
    _check_same(want, got)
    assert ms > 0
    small = cases[40:75]
    print(f"[no-alloc] call 2: {len(small)} extensions", flush=True)
    got2, _ = ga.extend_batch_noalloc(ctx, g, p, [c[0] for c in small], [c[1] for c in small], [c[2] for c in small])
    _check_same(want[40:75], got2)
    # and it still agrees with the oracle
    op = o.preset_params("vg")
    for (seq, node, off), (st, sc, tr) in zip(small, got2):
        w = o.extend(og, op, seq, node, off)
        assert w[0] == st and w[1] == sc and np.array_equal(w[2], tr)


def test_noalloc_rejects_batches_beyond_scratch(workdir):
    gfa, names, seqs = workload("mini", workdir)
    og = o.Graph(gfa)
    idx = o.Index(og, 15, 20, 0.001)
    cases = extension_cases(og, idx, seqs, 4, seed=3)
    ctx = ga.Context(0)
    g = ga.Graph(gfa)
    p = ga.params_preset("vg")
    ga.extend_batch(ctx, g, p, [cases[0][0]], [cases[0][1]], [cases[0][2]])
    # more extensions than the entry point takes
    many = [cases[0]] * 5000
    print("[no-alloc] too many extensions", flush=True)
    with pytest.raises(RuntimeError, match="GA_ESCRATCH"):
        ga.extend_batch_noalloc(ctx, g, p, [c[0] for c in many], [c[1] for c in many], [c[2] for c in many])
    # sequences and workspaces beyond the 128 MiB scratch
    # (two 1.5 Mbp extensions: 48 MB of traces and over 100 MB of workspaces)
    big = (cases[0][0] * 3000)[:1_500_000]
    assert len(big) == 1_500_000
    print("[no-alloc] too many bases", flush=True)
    with pytest.raises(RuntimeError, match="GA_ESCRATCH"):
        ga.extend_batch_noalloc(ctx, g, p, [big] * 2, [cases[0][1]] * 2, [cases[0][2]] * 2)
