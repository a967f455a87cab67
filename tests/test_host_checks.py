"""Host-side behaviour of the product pipeline (built for the host over the
device code run on CPU threads), compared with the oracle where the reference
fixes the result:

* an exception in a pipeline worker thread (a trace the replay cannot decode, a
  seed match outside the index) reaches the caller as an error and the process
  keeps running;
* the traceback's partial column recomputation changes nothing: a build that
  recomputes every column and keeps all of recalcNodeWordslice's checks
  (src/GraphAlignerBitvectorCommon.h:1078-1102) gives the same GAF and read
  statuses;
* the selection's --min-alignment-score passes through an int
  (src/AlignmentSelection.h:21, src/Aligner.cpp:402), so a fractional threshold
  selects as its truncation does (oracle and product), and a threshold the int
  cannot hold turns the filter off (src/AlignmentSelection.cpp:36).
"""
import pytest

import oracle_ffi as O
import pipe_ffi as P
import workloads as W


def _pipe(gfa, preset, variant=""):
    sd = O.PRESET_SEEDING[preset]
    pg = P.Graph(gfa, variant)
    return sd, pg, P.Index(pg, sd["k"], sd["w"], sd["discard"])


# fault 1 hits the trace codes of the host-driven rounds (the device seed loop
# returns raw trace items); fault 2 hits the seeding of either mode
@pytest.mark.parametrize("fault,mode", [(1, "rounds"), (1, "rounds_spec"), (2, "rounds"), (2, "device_loop")])
def test_worker_exception_reaches_caller(workdir, fault, mode):
    gfa, names, seqs = W.workload("mini", workdir)
    sd, pg, pi = _pipe(gfa, "vg")
    params = O.preset_params("vg")
    names, seqs = names[:40], seqs[:40]
    pg.L.pipe_set_fault(fault)
    try:
        with pytest.raises(RuntimeError):
            P.align_reads(pi, params, sd["density"], names, seqs, **P.MODES[mode])
    finally:
        pg.L.pipe_set_fault(0)
    # the library still works afterwards
    text, status, _ = P.align_reads(pi, params, sd["density"], names, seqs, **P.MODES[mode])
    og = O.Graph(gfa)
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    ot, ostat, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 4)
    assert text == ot and list(status) == list(ostat)


@pytest.mark.parametrize("wl,preset,n", [("mini", "vg", 0), ("mini_sv", "vg", 0), ("mini_sv", "dbg", 0), ("vg100k", "vg", 80)])
def test_full_recalculation_checks_change_nothing(workdir, wl, preset, n):
    gfa, names, seqs = W.workload(wl, workdir)
    if n:
        names, seqs = names[:n], seqs[:n]
    params = O.preset_params(preset)
    sd, pg, pi = _pipe(gfa, preset)
    _, fg, fi = _pipe(gfa, preset, "fullcheck")
    for mode in ("rounds_spec", "device_loop"):
        a = P.align_reads(pi, params, sd["density"], names, seqs, **P.MODES[mode])
        b = P.align_reads(fi, params, sd["density"], names, seqs, **P.MODES[mode])
        assert a[0] == b[0] and list(a[1]) == list(b[1])


def _as_values(text):
    return [float(f[5:]) for ln in text.split("\n") if ln for f in ln.split("\t") if f.startswith("AS:f:")]


@pytest.mark.parametrize("kind", ["fractional", "negative", "huge", "nan"])
def test_fractional_min_alignment_score(workdir, kind):
    import math
    gfa, names, seqs = W.workload("mini_sv", workdir)
    en, es = W.edge_reads(names, seqs)
    names, seqs = list(names[:60]) + en, list(seqs[:60]) + es
    sd, pg, pi = _pipe(gfa, "vg")
    og = O.Graph(gfa)
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    params = O.preset_params("vg")
    params.minAlignmentScore = 0.0
    base, _, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 4)
    scores = sorted(_as_values(base))
    # an alignment score x with a fractional part: threshold floor(x) + 0.99 lies
    # above x, but its int truncation floor(x) does not
    x = next(v for v in scores[len(scores) // 2:] if v - math.floor(v) < 0.9)
    threshold = {"fractional": math.floor(x) + 0.99, "negative": -3.7, "huge": 1e12, "nan": float("nan")}[kind]
    params.minAlignmentScore = threshold
    ot, ostat, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 4)
    pt, pstat, _ = P.align_reads(pi, params, sd["density"], names, seqs, **P.MODES["rounds_spec"])
    assert pt == ot and list(pstat) == list(ostat)
    kept = _as_values(ot)
    if kind == "fractional":
        assert x in kept and x < threshold
        assert min(kept) >= math.floor(x) and min(scores) < math.floor(x)
    elif kind in ("huge", "nan"):
        # the int cannot hold it: INT_MIN, so no score filter at all
        params.minAlignmentScore = -1e300
        nt, _, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 4)
        assert ot == nt and len(kept) >= len(scores)
    else:
        assert min(kept) >= -3


@pytest.mark.parametrize("preset", ["vg", "dbg"])
def test_traceback_through_self_loops(workdir, preset):
    """A traceback that re-enters the node it is in, in the same slice, through
    the node's edge to itself, at a column right of the columns it recomputed
    on entering: the columns are recomputed again (no stale column store)."""
    gfa, names, seqs = W.selfloop_workload(workdir)
    params = O.preset_params(preset)
    sd = O.PRESET_SEEDING[preset]
    og = O.Graph(gfa)
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    ot, ostat, _ = O.align_reads(og, oi, params, sd["density"], names, seqs, 4)
    assert (ostat == 0).sum() >= 20
    for variant in ("", "fullcheck"):
        _, pg, pi = _pipe(gfa, preset, variant)
        for mode in ("rounds", "device_loop"):
            pt, pstat, _ = P.align_reads(pi, params, sd["density"], names, seqs, **P.MODES[mode])
            assert list(pstat) == list(ostat), (variant, mode)
            assert pt == ot
