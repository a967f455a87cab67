"""The wave-cooperative lane code of the extension core (one extension per
wavefront: ballots for searches, shuffles for merges and X-scores, a node's
columns one per lane) on 64 host threads per wave that meet at a barrier in
every wave intrinsic, so that a lane leaving the others' control flow is
reported instead of waiting forever. The cases are the four extensions of the
mini workload on which the GPU kernel built from this lane code stopped
producing output (read lengths 513, 280, 2084, 526), one at a time with the
slice words checked against the oracle, and all four through the kernel's
persistent work loop (lane 0 takes the next request from a shared counter and
broadcasts it with a shuffle) with three waves."""
import ctypes
import os

import numpy as np
import pytest

import oracle_ffi as O
from emu_ffi import SliceSummary
from workloads import extension_cases, workload

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
i64p = ctypes.POINTER(ctypes.c_int64)


@pytest.fixture(scope="module")
def simt():
    L = ctypes.CDLL(os.path.join(ROOT, "tests", "_build", "libga_simt.so"))
    L.simt_graph_load.restype = ctypes.c_void_p
    L.simt_graph_load.argtypes = [ctypes.c_char_p]
    L.simt_graph_free.argtypes = [ctypes.c_void_p]
    L.simt_extend.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p,
                              ctypes.c_int64, ctypes.c_int, ctypes.c_int64, i64p, ctypes.c_int64, i64p,
                              ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(SliceSummary), ctypes.c_int64, i64p]
    L.simt_persistent.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                  ctypes.POINTER(ctypes.c_char_p), i64p, ctypes.POINTER(ctypes.c_int32), i64p, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    return L


@pytest.fixture(scope="module")
def hung_cases(workdir):
    gfa, names, seqs = workload("mini", workdir)
    sd = O.PRESET_SEEDING["vg"]
    og = O.Graph(gfa)
    oi = O.Index(og, sd["k"], sd["w"], sd["discard"])
    cases = extension_cases(og, oi, seqs, 4)
    assert [len(c[0]) for c in cases] == [513, 280, 2084, 526]
    return gfa, og, cases


@pytest.mark.parametrize("k", range(4))
def test_simt_lane_code_on_hung_cases(simt, hung_cases, k):
    gfa, og, cases = hung_cases
    seq, node, off = cases[k]
    p = O.preset_params("vg")
    g = simt.simt_graph_load(gfa.encode())
    try:
        cap = 4 * len(seq) + 1024
        trace = np.zeros(4 * cap, dtype=np.int64)
        nt, ns = ctypes.c_int64(), ctypes.c_int64()
        score = ctypes.c_int32()
        scap = len(seq) // 64 + 4
        sums = (SliceSummary * scap)()
        rc = simt.simt_extend(g, p.bandwidth, p.maxCellsPerSlice, p.errorCost, p.xdrop, seq.encode(), len(seq), node, off,
                              trace.ctypes.data_as(i64p), cap, ctypes.byref(nt), ctypes.byref(score), sums, scap, ctypes.byref(ns))
    finally:
        simt.simt_graph_free(g)
    assert rc >= 0, {-1: "lanes diverged", -2: "lanes disagree on the result", -3: "lanes disagree on the trace",
                     -4: "lanes disagree on the workspace"}.get(rc, rc)
    want = O.extend(og, p, seq, node, off)
    rows = [(s.j, s.minScore, s.minScoreNode, s.minScoreNodeOffset, s.maxExactEndposScore, s.maxExactEndposNode, s.cellsProcessed,
             s.numNodes, s.itemsHash, s.scoresNotValid) for s in sums[: ns.value]]
    assert rc == want[0] and score.value == want[1]
    assert np.array_equal(trace[: 4 * nt.value].reshape(-1, 4), want[2])
    assert rows == want[3]


def test_simt_persistent_work_loop(simt, hung_cases):
    gfa, og, cases = hung_cases
    p = O.preset_params("vg")
    n = len(cases)
    seqs = (ctypes.c_char_p * n)(*[c[0].encode() for c in cases])
    lens = np.array([len(c[0]) for c in cases], dtype=np.int64)
    nodes = np.array([c[1] for c in cases], dtype=np.int32)
    offs = np.array([c[2] for c in cases], dtype=np.int64)
    st = np.zeros(n, dtype=np.int32)
    sc = np.zeros(n, dtype=np.int32)
    g = simt.simt_graph_load(gfa.encode())
    try:
        rc = simt.simt_persistent(g, p.bandwidth, p.maxCellsPerSlice, p.errorCost, p.xdrop, n, seqs, lens.ctypes.data_as(i64p),
                                  nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), offs.ctypes.data_as(i64p), 3,
                                  st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), sc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    finally:
        simt.simt_graph_free(g)
    assert rc == 0
    for k, (seq, node, off) in enumerate(cases):
        want = O.extend(og, p, seq, node, off)
        assert st[k] == want[0] and sc[k] == want[1]
