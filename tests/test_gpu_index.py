"""The minimizer index built on the GPU (ga_index_build_gpu) against the host
build (pinned against the oracle in test_index_export.py): identical k-mers,
position ranges, positions in storage order, k-mer count and maxCount; and the
seeds found through either index identical."""
import os

import numpy as np
import pytest

import graphaligner_amd as ga
import workloads as W
from test_index_export import graphs, product_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0)


def _same_index(a, b):
    ea, eb = a.export(), b.export()
    for x, y, name in zip(ea, eb, ("kmers", "starts", "counts", "positions")):
        assert np.array_equal(x, y), name
    assert a.num_kmers == b.num_kmers
    assert a.max_count == b.max_count


@pytest.mark.parametrize("case", range(4))
def test_gpu_index_equals_host_index(ctx, workdir, case):
    gfa, preset, kw = graphs(workdir)[case]
    p = product_params(preset, kw)
    print(f"[gpu-index] case={case} {os.path.basename(gfa)} k={p.minimizerLength} w={p.minimizerWindowSize}", flush=True)
    g = ga.Graph(gfa)
    _same_index(ga.Index(g, p, ctx=ctx), ga.Index(g, p))


@pytest.mark.parametrize("wl,preset,kw", [("vg100k", "vg", None), ("vg5m", "vg", None), ("vg5m", "vg", (21, 31)),
                                          ("mini_sv", "vg", (15, 15)), ("mini_sv", "vg", (3, 40)), ("mini_dbg", "dbg", (31, 40))])
def test_gpu_index_equals_host_index_on_workloads(ctx, workdir, wl, preset, kw):
    gfa, names, seqs = W.workload(wl, workdir)
    p = product_params(preset, kw)
    print(f"[gpu-index] {wl} k={p.minimizerLength} w={p.minimizerWindowSize}", flush=True)
    g = ga.Graph(gfa)
    gi = ga.Index(g, p, ctx=ctx)
    hi = ga.Index(g, p)
    _same_index(gi, hi)
    # seeds of some reads through either index
    sample = seqs[:200]
    a = ga.seeds(ctx, g, gi, sample, p.minimizerSeedDensity)
    b = ga.seeds(ctx, g, hi, sample, p.minimizerSeedDensity)
    assert len(a) == len(b) and all(np.array_equal(x, y) for x, y in zip(a, b))
