"""ctypes bindings of the CPU oracle (oracle/_build/libga_oracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the parity checker, never by the product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(os.path.dirname(_HERE), "oracle", "_build")
# the reference's assertions on (as it ships) or compiled out (-DNDEBUG, its
# faster variant: used only to time the CPU baseline)
_LIB_PATHS = {False: os.path.join(_BUILD, "libga_oracle.so"), True: os.path.join(_BUILD, "libga_oracle_ndebug.so"),
              # with the reference's debug-only invariant checks restated and counted
              "checked": os.path.join(_BUILD, "libga_oracle_checked.so"),
              # the band queue's ties inside strongly connected components broken
              # the other way (tie-order sensitivity)
              "revties": os.path.join(_BUILD, "libga_oracle_revties.so")}


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("bandwidth", ctypes.c_int64),
        ("maxCellsPerSlice", ctypes.c_int64),
        ("errorCost", ctypes.c_int32),
        ("xdrop", ctypes.c_int32),
        ("seedExtendDensity", ctypes.c_double),
        ("multimapScoreFraction", ctypes.c_double),
        ("sloppy", ctypes.c_int32),
        ("minClusterSize", ctypes.c_int32),
        ("minAlignmentScore", ctypes.c_double),
        ("cigarMatchMismatchMerge", ctypes.c_int32),
    ]


class SliceSummary(ctypes.Structure):
    _fields_ = [
        ("j", ctypes.c_int64),
        ("minScore", ctypes.c_int32),
        ("minScoreNode", ctypes.c_int64),
        ("minScoreNodeOffset", ctypes.c_int64),
        ("maxExactEndposScore", ctypes.c_int32),
        ("maxExactEndposNode", ctypes.c_int64),
        ("cellsProcessed", ctypes.c_int64),
        ("numNodes", ctypes.c_int64),
        ("itemsHash", ctypes.c_uint64),
        ("scoresNotValid", ctypes.c_int32),
    ]


_libs = {}


def lib(ndebug=False):
    if ndebug not in _libs:
        path = _LIB_PATHS[ndebug]
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run make -C oracle")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.orc_graph_load.restype = vp
        L.orc_graph_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_graph_free.argtypes = [vp]
        L.orc_graph_num_nodes.restype = ctypes.c_int64
        L.orc_graph_num_nodes.argtypes = [vp]
        L.orc_graph_num_edges.restype = ctypes.c_int64
        L.orc_graph_num_edges.argtypes = [vp]
        L.orc_graph_first_ambiguous.restype = ctypes.c_int64
        L.orc_graph_first_ambiguous.argtypes = [vp]
        L.orc_graph_export.argtypes = [vp] + [ctypes.c_void_p] * 13
        L.orc_index_build.restype = vp
        L.orc_index_build.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_index_free.argtypes = [vp]
        L.orc_index_maxcount.restype = ctypes.c_int64
        L.orc_index_maxcount.argtypes = [vp]
        L.orc_index_num_kmers.restype = ctypes.c_int64
        L.orc_index_num_kmers.argtypes = [vp]
        L.orc_get_seeds.argtypes = [vp, vp, ctypes.c_char_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int, i64p, ctypes.c_int64, i64p]
        L.orc_extend.argtypes = [vp, ctypes.POINTER(OrcParams), ctypes.c_char_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int64,
                                 i64p, ctypes.c_int64, i64p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(SliceSummary), ctypes.c_int64, i64p]
        L.orc_extend_scheduled.argtypes = [vp, ctypes.POINTER(OrcParams), ctypes.c_char_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int64,
                                           ctypes.POINTER(ctypes.c_int32), ctypes.c_int64,
                                           i64p, ctypes.c_int64, i64p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(SliceSummary), ctypes.c_int64, i64p]
        L.orc_align_reads.restype = ctypes.c_double
        L.orc_align_reads.argtypes = [vp, vp, ctypes.POINTER(OrcParams), ctypes.c_double, ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p), i64p, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int32)]
        L.orc_free.argtypes = [vp]
        L.orc_order_seeds.restype = ctypes.c_int
        L.orc_order_seeds.argtypes = [vp, i64p, ctypes.c_int64]
        L.orc_align_one_way.restype = ctypes.c_int
        L.orc_align_one_way.argtypes = [vp, ctypes.POINTER(OrcParams), ctypes.c_char_p, ctypes.c_int64, i64p, ctypes.c_int64,
                                        ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        u64 = ctypes.c_uint64
        L.orc_merge.argtypes = [u64, u64, ctypes.c_int32, u64, u64, ctypes.c_int32, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_int32)]
        L.orc_difference_masks.argtypes = [u64, u64, u64, u64, ctypes.c_int32, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.orc_max_xscore.restype = ctypes.c_int32
        L.orc_max_xscore.argtypes = [u64, u64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, u64]
        L.orc_changed_min_score.restype = ctypes.c_int32
        L.orc_changed_min_score.argtypes = [u64, u64, ctypes.c_int32, u64, u64, ctypes.c_int32]
        L.orc_next_slice.argtypes = [u64, u64, u64, ctypes.c_int32, u64, u64, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                     ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.orc_last_word_steps.restype = ctypes.c_int64
        L.orc_minimizers.restype = ctypes.c_int64
        L.orc_minimizers.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, i64p, ctypes.c_int64]
        L.orc_index_table.restype = ctypes.c_int64
        L.orc_index_table.argtypes = [vp, vp, vp, vp]
        L.orc_extra_checks_run.restype = ctypes.c_int64
        L.orc_extra_check_failures.restype = ctypes.c_int64
        L.orc_slice_summary_size.restype = ctypes.c_int64
        assert L.orc_slice_summary_size() == ctypes.sizeof(SliceSummary)
        _libs[ndebug] = L
    return _libs[ndebug]


def preset_params(preset="vg"):
    """Preset values of src/AlignerMain.cpp:157-183 and the X-drop default :220-231."""
    p = OrcParams()
    p.errorCost = 294
    p.xdrop = 14705
    p.multimapScoreFraction = 0.9
    p.sloppy = 1
    p.minClusterSize = 1
    if preset == "vg":
        p.bandwidth = 10
        p.maxCellsPerSlice = 0
        p.seedExtendDensity = -1.0
    elif preset == "dbg":
        p.bandwidth = 5
        p.maxCellsPerSlice = 10000
        p.seedExtendDensity = 0.002
    else:
        raise ValueError(preset)
    return p


PRESET_SEEDING = {"vg": dict(k=15, w=20, density=10.0, discard=0.001),
                  "dbg": dict(k=19, w=30, density=5.0, discard=0.0002)}


class Graph:
    def __init__(self, path, ndebug=False):
        self.L = lib(ndebug)
        err = ctypes.create_string_buffer(512)
        self.h = self.L.orc_graph_load(path.encode(), err, 512)
        if not self.h:
            raise RuntimeError("oracle graph load failed: " + err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_graph_free(self.h)
            self.h = None

    def export(self):
        L = self.L
        n = L.orc_graph_num_nodes(self.h)
        e = L.orc_graph_num_edges(self.h)
        a = {k: np.zeros(n, dtype=np.int64) for k in ["nodeLength", "nodeIDs", "nodeOffset", "comp", "chain", "chainPos"]}
        lin = np.zeros(n, dtype=np.uint8)
        rev = np.zeros(n, dtype=np.uint8)
        seq2 = np.zeros(2 * n, dtype=np.uint64)
        inOff = np.zeros(n + 1, dtype=np.int64)
        outOff = np.zeros(n + 1, dtype=np.int64)
        inAdj = np.zeros(max(e, 1), dtype=np.int64)
        outAdj = np.zeros(max(e, 1), dtype=np.int64)
        P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        L.orc_graph_export(self.h, P(a["nodeLength"]), P(a["nodeIDs"]), P(a["nodeOffset"]), P(a["comp"]), P(lin), P(rev),
                           P(a["chain"]), P(a["chainPos"]), P(seq2), P(inOff), P(inAdj), P(outOff), P(outAdj))
        a.update(linearizable=lin, reverse=rev, seq2=seq2, inOff=inOff, inAdj=inAdj[:e], outOff=outOff, outAdj=outAdj[:e],
                 firstAmbiguous=L.orc_graph_first_ambiguous(self.h))
        return a


class Index:
    def __init__(self, graph, k, w, discard):
        self.graph = graph
        self.L = graph.L
        self.h = self.L.orc_index_build(graph.h, k, w, discard)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_index_free(self.h)
            self.h = None

    @property
    def max_count(self):
        return self.L.orc_index_maxcount(self.h)

    def seeds(self, seq, density, order=False, cap_per_base=64):
        cap = max(1024, cap_per_base * len(seq))
        out = np.zeros(10 * cap, dtype=np.int64)
        n = ctypes.c_int64()
        rc = self.L.orc_get_seeds(self.graph.h, self.h, seq.encode(), len(seq), density, int(order),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), cap, ctypes.byref(n))
        if rc == 2:
            return None  # OrderSeeds fired a reference assertion: the read is dropped
        assert rc == 0
        return out[: 10 * n.value].reshape(-1, 10)


def extend(graph, params, seq, bigraph_node, offset, schedule=None):
    """getReverseTraceFromSeed: returns (rc, score, trace[n,4], summaries list).

    schedule: optional per-slice bandwidths (slice s uses schedule[min(s, len-1)])."""
    cap = 4 * len(seq) + 1024
    trace = np.zeros(4 * cap, dtype=np.int64)
    nt = ctypes.c_int64()
    score = ctypes.c_int32()
    scap = len(seq) // 64 + 4
    sums = (SliceSummary * scap)()
    ns = ctypes.c_int64()
    tp = trace.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    if schedule is None or len(schedule) == 0:
        rc = graph.L.orc_extend(graph.h, ctypes.byref(params), seq.encode(), len(seq), bigraph_node, offset,
                              tp, cap, ctypes.byref(nt), ctypes.byref(score), sums, scap, ctypes.byref(ns))
    else:
        sch = np.ascontiguousarray(schedule, dtype=np.int32)
        rc = graph.L.orc_extend_scheduled(graph.h, ctypes.byref(params), seq.encode(), len(seq), bigraph_node, offset,
                                        sch.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(sch),
                                        tp, cap, ctypes.byref(nt), ctypes.byref(score), sums, scap, ctypes.byref(ns))
    rows = [(s.j, s.minScore, s.minScoreNode, s.minScoreNodeOffset, s.maxExactEndposScore, s.maxExactEndposNode,
             s.cellsProcessed, s.numNodes, s.itemsHash, s.scoresNotValid) for s in sums[: ns.value]]
    return rc, score.value, trace[: 4 * nt.value].reshape(-1, 4), rows


def align_reads(graph, index, params, density, names, seqs, nthreads=1):
    n = len(seqs)
    names_b = [x.encode() for x in names]
    seqs_b = [x.encode() for x in seqs]
    na = (ctypes.c_char_p * n)(*names_b)
    sa = (ctypes.c_char_p * n)(*seqs_b)
    lens = np.array([len(s) for s in seqs], dtype=np.int64)
    out = ctypes.c_void_p()
    status = np.zeros(n, dtype=np.int32)
    secs = graph.L.orc_align_reads(graph.h, index.h, ctypes.byref(params), density, n, na, sa,
                                 lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), nthreads, ctypes.byref(out),
                                 status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    text = ctypes.string_at(out.value).decode()
    graph.L.orc_free(out)
    return text, status, secs


def align_one_way(graph, params, seq, seeds, select=False, mapq=False):
    """GraphAligner::AlignOneWay with the given seeds (int64 rows of 10 SeedHit
    fields, -1 for an unset alignment-graph node), optionally followed by
    AddMappingQualities and the selection/output sorts. Returns text: "S status
    seedsExtended", then per alignment an "A ..." line and a trace line."""
    seeds = np.ascontiguousarray(seeds, dtype=np.int64).reshape(-1, 10)
    out = ctypes.c_void_p()
    graph.L.orc_align_one_way(graph.h, ctypes.byref(params), seq.encode(), len(seq),
                              seeds.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(seeds), int(select), int(mapq),
                              ctypes.byref(out))
    text = ctypes.string_at(out.value).decode()
    graph.L.orc_free(out)
    return text


def order_seeds(graph, seeds):
    """OrderSeeds on a copy of the seed rows; None when a reference assertion fires."""
    out = np.ascontiguousarray(seeds, dtype=np.int64).reshape(-1, 10).copy()
    rc = graph.L.orc_order_seeds(graph.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(out))
    return None if rc == 2 else out


def minimizers(seq, k, w, L=None):
    """iterateMinimizersReal over one string: int64 array [n, 2] of (position, k-mer)."""
    L = L or lib()
    cap = 2 * len(seq) + 16
    out = np.zeros(2 * cap, dtype=np.int64)
    n = L.orc_minimizers(seq.encode(), len(seq), k, w, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), cap)
    assert n <= cap
    return out[: 2 * n].reshape(-1, 2)


def index_table(index):
    """{k-mer: positions in storage order} of an oracle index."""
    L = index.L
    total = L.orc_index_table(index.h, None, None, None)
    nk = L.orc_index_num_kmers(index.h)
    kmers = np.zeros(max(nk, 1), dtype=np.uint64)
    counts = np.zeros(max(nk, 1), dtype=np.int64)
    pos = np.zeros(max(total, 1), dtype=np.uint64)
    L.orc_index_table(index.h, kmers.ctypes.data_as(ctypes.c_void_p), counts.ctypes.data_as(ctypes.c_void_p), pos.ctypes.data_as(ctypes.c_void_p))
    out, at = {}, 0
    for km, c in zip(kmers[:nk].tolist(), counts[:nk].tolist()):
        out[km] = pos[at: at + c].tolist()
        at += c
    return out
