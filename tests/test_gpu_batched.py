"""The batched persistent form on gfx950 (deviceSeedLoop = 2: a fixed set of
workers pulling whole reads from a device counter, one captured HIP graph per
batch, worker count autotuned) against the plain GPU path and the oracle, on
the mixed read stream of the 5 Mbp variation graph: log-normal lengths (median
10 kbp, sigma 0.9, clipped to 1-100 kbp), per-read error 5-15 %, minimizers of
21 bases in windows of 31 (5th-95th percentile 5-193 seeds per read). The graph
is replayed batch after batch, with other GPU work in between, and every replay
is checked."""
import os

import numpy as np
import pytest

import oracle_ffi as o
import graphaligner_amd as ga
from graphaligner_amd import synth

pytestmark = pytest.mark.gpu

STREAM_K, STREAM_W = 21, 31


@pytest.fixture(scope="module")
def stream(workdir):
    prefix = os.path.join(workdir, "stream")
    synth.generate("vg5m_stream", prefix, overrides_reads=dict(numReads=2400))
    names, seqs = synth.read_fasta(prefix + ".fa")
    return prefix + ".gfa", names, seqs


def _params(mode):
    p = ga.params_preset("vg")
    p.minimizerLength, p.minimizerWindowSize = STREAM_K, STREAM_W
    p.deviceSeedLoop = mode
    return p


def test_gpu_batched_matches_plain_and_oracle(stream):
    gfa, names, seqs = stream
    ctx = ga.Context(0)
    g = ga.Graph(gfa)
    gi = ga.Index(g, _params(0))
    og = o.Graph(gfa)
    oi = o.Index(og, STREAM_K, STREAM_W, 0.001)
    lens = np.array([len(s) for s in seqs])
    print(f"[gpu-batched] {len(seqs)} reads, {lens.sum() / 1e6:.1f} Mbp, longest {lens.max()}", flush=True)
    ot, ostat, _ = o.align_reads(og, oi, o.preset_params("vg"), 10.0, names, seqs, 16)
    plain, pstat, _ = ga.align_reads(ctx, g, gi, _params(0), names, seqs)
    assert plain == ot and list(pstat) == list(ostat)
    # batches of 600 reads: one graph, replayed, with plain-path work between
    # the replays; every replay checked against the plain path
    workers = None
    for b in range(0, len(seqs), 600):
        bn, bs = names[b:b + 600], seqs[b:b + 600]
        print(f"[gpu-batched] batch {b // 600}", flush=True)
        text, st, stats = ga.align_reads(ctx, g, gi, _params(2), bn, bs)
        want, wst, _ = ga.align_reads(ctx, g, gi, _params(1), bn, bs)
        assert text == want and list(st) == list(wst)
        assert list(st) == list(ostat[b:b + 600])
        assert stats["persistentWorkers"] > 0
        workers = workers or stats["persistentWorkers"]
        assert stats["persistentWorkers"] == workers  # tuned once per context
    lines = [ln for ln in ot.split("\n") if ln]
    assert len(lines) > len(seqs)
