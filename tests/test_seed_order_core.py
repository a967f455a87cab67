"""Seed expansion and OrderSeeds as the GPU runs them (ga_order_core.h), built
for the host: libstdSort gives libstdc++'s std::sort permutation (ties
included, heap-sort fallback too), and the seeds of whole reads equal the host
pipeline's (seedsFromMatches + orderSeedsChecked, themselves equal to the
oracle's) on variation graphs and the de Bruijn graph, edge reads included."""
import pytest

import pipe_ffi as P
import oracle_ffi as O
import workloads as W


@pytest.mark.parametrize("n", [0, 1, 2, 3, 15, 16, 17, 31, 33, 100, 257, 1000, 5000])
@pytest.mark.parametrize("ties", [1, 2, 5, 50, 1 << 40])
def test_libstd_sort_permutation(n, ties):
    L = P.lib()
    for seed in range(4):
        assert L.pipe_sort_check(n, seed * 1000 + n, ties, 0) == 0
        assert L.pipe_sort_check(n, seed * 1000 + n, ties, 1) == 0


@pytest.mark.parametrize("wl,preset", [("mini_sv", "vg"), ("vg100k", "vg"), ("mini_dbg", "dbg")])
def test_device_seeding_code_equals_host_seeds(workdir, wl, preset):
    gfa, names, seqs = W.workload(wl, workdir)
    sd = O.PRESET_SEEDING[preset]
    g = P.Graph(gfa)
    idx = P.Index(g, sd["k"], sd["w"], sd["discard"])
    reads = seqs[:150] + W.edge_reads(names, seqs)[1]
    compared = 0
    for s in reads:
        want = idx.seeds(s, sd["density"], order=True)
        got = idx.seeds_device_code(s, sd["density"])
        if want is None or got is None:
            assert want is None and got is None
            continue
        assert (want == got).all() if len(want) else len(got) == 0
        compared += len(want)
    assert compared > 1000
