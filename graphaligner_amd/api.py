"""Python interface over the C ABI, shaped like the reference's wrapper functions.

* Graph          ~ AlignmentGraph from GfaGraph::LoadFromFile + DirectedGraph::BuildFromGFA
* extend_batch   ~ GraphAlignerBitvectorBanded::getReverseTraceFromSeed, batched
* Index          ~ MinimizerSeeder
* order_seeds    ~ OrderSeeds (src/GraphAlignerWrapper.h:59)
* align_one_way  ~ AlignOneWay with the caller's seeds (src/GraphAlignerWrapper.h:55),
                   batched over reads; returns an AlignmentResults object with
                   select / add_mapping_qualities / gaf_line (AddGAFLine)
* align_reads    ~ the per-read pipeline of src/Aligner.cpp:390-606 with GAF output
"""
import ctypes
import weakref

import numpy as np

from . import _native as N


def _check(rc, what, err=None):
    if rc != N.GA_OK:
        msg = N.STATUS_NAMES.get(rc, str(rc))
        if err is not None and err.value:
            msg += ": " + err.value.decode(errors="replace")
        raise RuntimeError(f"{what} failed with {msg}")


def params_preset(preset="vg"):
    p = N.Params()
    _check(N.lib().ga_params_preset(preset.encode() if preset else None, ctypes.byref(p)), "ga_params_preset")
    return p


class Graph:
    def __init__(self, path):
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        _check(N.lib().ga_graph_load_gfa(path.encode(), ctypes.byref(h), err, 1024), "ga_graph_load_gfa(" + path + ")", err)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().ga_graph_free(self.h)
            self.h = None

    @property
    def num_nodes(self):
        return N.lib().ga_graph_num_nodes(self.h)

    def export(self):
        L = N.lib()
        n = L.ga_graph_num_nodes(self.h)
        e = L.ga_graph_num_edges(self.h)
        a = {k: np.zeros(n, dtype=np.int64) for k in ["nodeLength", "nodeIDs", "nodeOffset", "comp", "chain", "chainPos"]}
        lin = np.zeros(n, dtype=np.uint8)
        rev = np.zeros(n, dtype=np.uint8)
        seq2 = np.zeros(2 * n, dtype=np.uint64)
        inOff = np.zeros(n + 1, dtype=np.int64)
        outOff = np.zeros(n + 1, dtype=np.int64)
        inAdj = np.zeros(max(e, 1), dtype=np.int64)
        outAdj = np.zeros(max(e, 1), dtype=np.int64)
        P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        _check(L.ga_graph_export(self.h, P(a["nodeLength"]), P(a["nodeIDs"]), P(a["nodeOffset"]), P(a["comp"]), P(lin), P(rev),
                                 P(a["chain"]), P(a["chainPos"]), P(seq2), P(inOff), P(inAdj), P(outOff), P(outAdj)), "ga_graph_export")
        a.update(linearizable=lin, reverse=rev, seq2=seq2, inOff=inOff, inAdj=inAdj[:e], outOff=outOff, outAdj=outAdj[:e],
                 firstAmbiguous=L.ga_graph_first_ambiguous(self.h))
        return a


class Context:
    """One GPU. Raises if no HIP device is present: there is no CPU path."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        _check(N.lib().ga_context_create(device, ctypes.byref(h), err, 1024), "ga_context_create", err)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().ga_context_free(self.h)
            self.h = None

    def synchronize(self):
        _check(N.lib().ga_context_synchronize(self.h), "ga_context_synchronize")


def extend_batch(ctx, graph, params, seqs, bigraph_nodes, offsets, summaries=False, schedules=None):
    """Returns per extension (status, score, trace[n,4] end first, summaries).

    schedules: optional list of per-slice bandwidth sequences, one per extension
    (slice s uses schedule[min(s, len - 1)]; an empty one keeps params.bandwidth),
    run through ga_extend_batch_scheduled."""
    n = len(seqs)
    seqs_b = [s.encode() for s in seqs]
    lens = np.array([len(s) for s in seqs], dtype=np.int64)
    cap = int(2 * lens.max() + 256) if n else 256
    traces = np.zeros(4 * cap * max(n, 1), dtype=np.int64)
    ntrace = np.zeros(n, dtype=np.int64)
    scores = np.zeros(n, dtype=np.int32)
    status = np.zeros(n, dtype=np.int32)
    nsl = np.zeros(n, dtype=np.int64)
    scap = int(lens.max() // 64 + 4) if n else 1
    sums = (N.SliceSummary * (scap * n))() if summaries else None
    ms = ctypes.c_double()
    arr = (ctypes.c_char_p * n)(*seqs_b)
    nodes = np.asarray(bigraph_nodes, dtype=np.int32)
    offs = np.asarray(offsets, dtype=np.int64)
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    if schedules is None:
        rc = N.lib().ga_extend_batch(ctx.h, graph.h, ctypes.byref(params), n, arr, P(lens), P(nodes), P(offs), cap, P(traces),
                                     P(ntrace), P(scores), P(status), sums, scap, P(nsl), ctypes.byref(ms))
        _check(rc, "ga_extend_batch")
    else:
        if len(schedules) != n:
            raise ValueError("one schedule per extension")
        slens = np.array([len(x) for x in schedules], dtype=np.int64)
        soffs = np.concatenate([[0], np.cumsum(slens)[:-1]]).astype(np.int64) if n else np.zeros(0, dtype=np.int64)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.int32) for x in schedules] + [np.zeros(1, np.int32)]))
        rc = N.lib().ga_extend_batch_scheduled(ctx.h, graph.h, ctypes.byref(params), n, arr, P(lens), P(nodes), P(offs), P(flat),
                                               P(soffs), P(slens), cap, P(traces), P(ntrace), P(scores), P(status), sums, scap,
                                               P(nsl), ctypes.byref(ms))
        _check(rc, "ga_extend_batch_scheduled")
    out = []
    for i in range(n):
        t = traces[4 * cap * i: 4 * cap * i + 4 * ntrace[i]].reshape(-1, 4)
        rows = []
        if summaries:
            for s in sums[i * scap: i * scap + nsl[i]]:
                rows.append((s.j, s.minScore, s.minScoreNode, s.minScoreNodeOffset, s.maxExactEndposScore, s.maxExactEndposNode,
                             s.cellsProcessed, s.numNodes, s.itemsDigest, s.scoresNotValid))
        out.append((int(status[i]), int(scores[i]), t, rows))
    return out, ms.value


def extend_batch_noalloc(ctx, graph, params, seqs, bigraph_nodes, offsets):
    """extend_batch through the entry point that allocates no device memory
    (ga_extend_batch_noalloc; the graph must be resident on the context).
    Returns per extension (status, score, trace[n,4] end first) and the kernel ms."""
    n = len(seqs)
    seqs_b = [s.encode() for s in seqs]
    lens = np.array([len(s) for s in seqs], dtype=np.int64)
    cap = int(2 * lens.max() + 256) if n else 256
    traces = np.zeros(4 * cap * max(n, 1), dtype=np.int64)
    ntrace = np.zeros(max(n, 1), dtype=np.int64)
    scores = np.zeros(max(n, 1), dtype=np.int32)
    status = np.zeros(max(n, 1), dtype=np.int32)
    ms = ctypes.c_double()
    arr = (ctypes.c_char_p * max(n, 1))(*seqs_b)
    nodes = np.asarray(bigraph_nodes, dtype=np.int32)
    offs = np.asarray(offsets, dtype=np.int64)
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    _check(N.lib().ga_extend_batch_noalloc(ctx.h, graph.h, ctypes.byref(params), n, arr, P(lens), P(nodes), P(offs), cap,
                                           P(traces), P(ntrace), P(scores), P(status), ctypes.byref(ms)), "ga_extend_batch_noalloc")
    out = [(int(status[i]), int(scores[i]), traces[4 * cap * i: 4 * cap * i + 4 * ntrace[i]].reshape(-1, 4)) for i in range(n)]
    return out, ms.value


class Index:
    """Minimizer index (MinimizerSeeder) built from params' minimizer settings,
    on the host, or on the context's GPU when ctx is given (ga_index_build_gpu)."""

    def __init__(self, graph, params, ctx=None):
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        if ctx is None:
            _check(N.lib().ga_index_build(graph.h, ctypes.byref(params), ctypes.byref(h), err, 1024), "ga_index_build", err)
        else:
            _check(N.lib().ga_index_build_gpu(ctx.h, graph.h, ctypes.byref(params), ctypes.byref(h), err, 1024), "ga_index_build_gpu", err)
        self.h = h
        self.graph = graph

    def export(self):
        """(k-mers ascending, starts, counts, positions) as numpy arrays."""
        n = self.num_kmers
        m = N.lib().ga_index_num_positions(self.h)
        kmers = np.zeros(max(n, 1), dtype=np.uint64)
        starts = np.zeros(max(n, 1), dtype=np.int64)
        counts = np.zeros(max(n, 1), dtype=np.int64)
        pos = np.zeros(max(m, 1), dtype=np.uint64)
        P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        _check(N.lib().ga_index_export(self.h, P(kmers), P(starts), P(counts), P(pos)), "ga_index_export")
        return kmers[:n], starts[:n], counts[:n], pos[:m]

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().ga_index_free(self.h)
            self.h = None

    @property
    def max_count(self):
        return N.lib().ga_index_max_count(self.h)

    @property
    def num_kmers(self):
        return N.lib().ga_index_num_kmers(self.h)


def _reads(seqs):
    n = len(seqs)
    b = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    arr = (ctypes.c_char_p * max(n, 1))(*b)
    lens = np.array([len(x) for x in b], dtype=np.int64)
    return n, arr, lens


def seeds(ctx, graph, index, seqs, density, order=True):
    """Per read an int64 array [nseeds, 10] in SeedHit field order."""
    n, arr, lens = _reads(seqs)
    counts = np.zeros(max(n, 1), dtype=np.int64)
    out = ctypes.c_void_p()
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    _check(N.lib().ga_seeds(ctx.h, graph.h, index.h, density, int(order), n, arr, P(lens), ctypes.byref(out), P(counts)), "ga_seeds")
    total = int(counts[:n].sum())
    flat = np.ctypeslib.as_array(ctypes.cast(out, ctypes.POINTER(ctypes.c_int64)), shape=(max(10 * total, 1),)).copy()
    N.lib().ga_free(out)
    res, pos = [], 0
    for i in range(n):
        c = int(counts[i])
        res.append(flat[10 * pos: 10 * (pos + c)].reshape(-1, 10))
        pos += c
    return res


_decode_utf8 = ctypes.pythonapi.PyUnicode_DecodeUTF8
_decode_utf8.restype = ctypes.py_object
_decode_utf8.argtypes = [ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_char_p]


def _gaf_text(out, st, gaf_format="str"):
    """The library's GAF buffer as str (decoded in place, one copy, then freed), or
    with gaf_format="buffer" as a uint8 array over the library's buffer itself (no
    copy; the buffer is freed when the array and its views are collected)."""
    if gaf_format == "buffer":
        n = int(st.gafBytes)
        if n == 0 or not out.value:
            if out.value:
                N.lib().ga_free(out)
            return np.zeros(0, dtype=np.uint8)
        raw = (ctypes.c_uint8 * n).from_address(out.value)
        weakref.finalize(raw, N.lib().ga_free, ctypes.c_void_p(out.value))
        return np.ctypeslib.as_array(raw)
    if gaf_format != "str":
        raise ValueError("gaf_format must be 'str' or 'buffer'")
    try:
        return _decode_utf8(out.value, st.gafBytes, b"strict")
    finally:
        N.lib().ga_free(out)


def align_reads(ctx, graph, index, params, names, seqs, gaf_format="str"):
    """GAF text of all reads (lines of read 0 first), per-read status, stats dict.
    gaf_format: "str", or "buffer" for the text's bytes without a copy (_gaf_text)."""
    n, arr, lens = _reads(seqs)
    na = (ctypes.c_char_p * max(n, 1))(*[x.encode() for x in names])
    out = ctypes.c_void_p()
    status = np.zeros(max(n, 1), dtype=np.int32)
    st = N.AlignStats()
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    _check(N.lib().ga_align_reads(ctx.h, graph.h, index.h, ctypes.byref(params), n, na, arr, P(lens), ctypes.byref(out),
                                  P(status), ctypes.byref(st)), "ga_align_reads")
    text = _gaf_text(out, st, gaf_format)
    stats = {f: getattr(st, f) for f, _ in N.AlignStats._fields_}
    return text, status[:n], stats


def upload_reads(ctx, seqs):
    """Makes reads resident on the context's GPU (replacing the previous set)."""
    n, arr, lens = _reads(seqs)
    _check(N.lib().ga_reads_upload(ctx.h, n, arr, lens.ctypes.data_as(ctypes.c_void_p)), "ga_reads_upload")


def align_resident(ctx, graph, index, params, first, names, gaf_format="str"):
    """align_reads on resident reads [first, first + len(names))."""
    count = len(names)
    na = (ctypes.c_char_p * max(count, 1))(*[x.encode() for x in names])
    out = ctypes.c_void_p()
    status = np.zeros(max(count, 1), dtype=np.int32)
    st = N.AlignStats()
    _check(N.lib().ga_align_resident(ctx.h, graph.h, index.h, ctypes.byref(params), first, count, na, ctypes.byref(out),
                                     status.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st)), "ga_align_resident")
    text = _gaf_text(out, st, gaf_format)
    stats = {f: getattr(st, f) for f, _ in N.AlignStats._fields_}
    return text, status[:count], stats


SEED_FIELDS = [f for f, _ in N.SeedHit._fields_]


def _seed_array(rows):
    rows = np.ascontiguousarray(rows, dtype=np.int64).reshape(-1, 10)
    arr = (N.SeedHit * max(len(rows), 1))()
    ctypes.memmove(arr, rows.ctypes.data, rows.nbytes)
    return arr, len(rows)


def order_seeds(graph, seeds):
    """OrderSeeds on a copy of the seed rows (int64 [n, 10] in SeedHit field
    order, -1 for an unset alignment-graph node); raises on a reference
    assertion (GA_ASSERT: the reference drops the read)."""
    arr, n = _seed_array(seeds)
    _check(N.lib().ga_order_seeds(graph.h, arr, n), "ga_order_seeds")
    return np.ctypeslib.as_array(ctypes.cast(arr, ctypes.POINTER(ctypes.c_int64)), shape=(max(n, 1) * 10,))[: 10 * n].reshape(-1, 10).copy()


class AlignmentResults:
    """AlignmentResult of each read of an align_one_way call (ga_alignment_result_t)."""

    def __init__(self, h):
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().ga_result_free(self.h)
            self.h = None

    def __len__(self):
        return N.lib().ga_result_reads(self.h)

    def status(self, read):
        return N.lib().ga_result_status(self.h, read)

    def seeds_extended(self, read):
        return N.lib().ga_result_seeds_extended(self.h, read)

    def count(self, read):
        return N.lib().ga_result_count(self.h, read)

    def item(self, read, i):
        it = N.AlignmentItem()
        _check(N.lib().ga_result_item(self.h, read, i, ctypes.byref(it)), "ga_result_item")
        return {f: getattr(it, f) for f, _ in N.AlignmentItem._fields_ if f != "pad"}

    def trace(self, read, i):
        """[(node, nodeOffset, seqPos, nodeSwitch, sequenceCharacter, graphCharacter)]"""
        n = self.item(read, i)["traceLength"]
        arr = (N.TraceItem * max(n, 1))()
        _check(N.lib().ga_result_trace(self.h, read, i, arr, n), "ga_result_trace")
        return [(t.node, t.nodeOffset, t.seqPos, int(t.nodeSwitch), t.sequenceCharacter, t.graphCharacter) for t in arr[:n]]

    def select(self, read, params, sort_for_output=True):
        _check(N.lib().ga_select_alignments(self.h, read, ctypes.byref(params), int(sort_for_output)), "ga_select_alignments")

    def add_mapping_qualities(self, read):
        _check(N.lib().ga_add_mapping_qualities(self.h, read), "ga_add_mapping_qualities")

    def gaf_line(self, graph, read, i, name, read_length, cigar_match_mismatch_merge=False):
        need = ctypes.c_int64()
        rc = N.lib().ga_gaf_line(graph.h, self.h, read, i, name.encode(), read_length, int(cigar_match_mismatch_merge), None, 0,
                                 ctypes.byref(need))
        if rc not in (N.GA_OK, N.GA_EBADARG) or need.value <= 0:
            _check(rc, "ga_gaf_line")
        buf = ctypes.create_string_buffer(need.value)
        _check(N.lib().ga_gaf_line(graph.h, self.h, read, i, name.encode(), read_length, int(cigar_match_mismatch_merge), buf,
                                   need.value, ctypes.byref(need)), "ga_gaf_line")
        return buf.value.decode()


def align_one_way(ctx, graph, params, seqs, seeds):
    """AlignOneWay for each read with its ordered seeds (seeds[i]: int64 [n_i, 10])."""
    n, arr, lens = _reads(seqs)
    counts = np.array([len(np.asarray(s).reshape(-1, 10)) for s in seeds], dtype=np.int64)
    flat = np.concatenate([np.asarray(s, dtype=np.int64).reshape(-1, 10) for s in seeds]) if n else np.zeros((0, 10), np.int64)
    sarr, _ = _seed_array(flat)
    h = ctypes.c_void_p()
    _check(N.lib().ga_align_one_way(ctx.h, graph.h, ctypes.byref(params), n, arr, lens.ctypes.data_as(ctypes.c_void_p), sarr,
                                     counts.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)), "ga_align_one_way")
    return AlignmentResults(h)


MLP_FEATURES, MLP_OUTPUTS = 48, 2


def mlp_num_params():
    return N.lib().ga_mlp_num_params()


def _f32(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(-1, cols)


class Mlp:
    """The learned proposal's network (48 -> 128 -> 128 -> 2, BF16 MFMA kernels on
    the context's GPU). Parameters are a flat float32 array of mlp_num_params()."""

    def __init__(self, ctx, seed=1):
        h = ctypes.c_void_p()
        _check(N.lib().ga_mlp_create(ctx.h, seed, ctypes.byref(h)), "ga_mlp_create")
        self.h = h
        self.ctx = ctx

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().ga_mlp_free(self.h)
            self.h = None

    def params(self):
        out = np.zeros(mlp_num_params(), dtype=np.float32)
        _check(N.lib().ga_mlp_get_params(self.h, out.ctypes.data_as(ctypes.c_void_p)), "ga_mlp_get_params")
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == mlp_num_params()
        _check(N.lib().ga_mlp_set_params(self.h, w.ctypes.data_as(ctypes.c_void_p)), "ga_mlp_set_params")

    def forward(self, features):
        x = _f32(features, MLP_FEATURES)
        out = np.zeros((len(x), MLP_OUTPUTS), dtype=np.float32)
        _check(N.lib().ga_mlp_forward(self.h, len(x), x.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p)),
               "ga_mlp_forward")
        return out

    def gradients(self, features, labels):
        """(mean binary cross-entropy, flat gradient)"""
        x = _f32(features, MLP_FEATURES)
        y = _f32(labels, MLP_OUTPUTS)
        assert len(x) == len(y)
        g = np.zeros(mlp_num_params(), dtype=np.float32)
        loss = ctypes.c_double()
        _check(N.lib().ga_mlp_gradients(self.h, len(x), x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.byref(loss), g.ctypes.data_as(ctypes.c_void_p)), "ga_mlp_gradients")
        return loss.value, g

    def train(self, features, labels, epochs=4, batch=4096, lr=2e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=1):
        """Adam on the GPU; returns (per-epoch mean loss, device ms)."""
        x = _f32(features, MLP_FEATURES)
        y = _f32(labels, MLP_OUTPUTS)
        assert len(x) == len(y)
        p = N.MlpTrainParams(epochs, batch, lr, beta1, beta2, eps, seed)
        losses = np.zeros(epochs, dtype=np.float32)
        ms = ctypes.c_double()
        _check(N.lib().ga_mlp_train(self.h, len(x), x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.byref(p), losses.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ms)), "ga_mlp_train")
        return losses, ms.value


def cluster_records(ctx, graph, index, params, first, count):
    """Outcome records of the plain method on resident reads [first, first + count):
    (features [n, 48], labels [n, 2], read index [n])."""
    h = ctypes.c_void_p()
    L = N.lib()
    _check(L.ga_cluster_records(ctx.h, graph.h, index.h, ctypes.byref(params), first, count, ctypes.byref(h)), "ga_cluster_records")
    try:
        n = L.ga_records_count(h)
        if n == 0:
            return np.zeros((0, MLP_FEATURES), np.float32), np.zeros((0, MLP_OUTPUTS), np.float32), np.zeros(0, np.int64)
        f = np.ctypeslib.as_array(ctypes.cast(L.ga_records_features(h), ctypes.POINTER(ctypes.c_float)), shape=(n, MLP_FEATURES)).copy()
        y = np.ctypeslib.as_array(ctypes.cast(L.ga_records_labels(h), ctypes.POINTER(ctypes.c_float)), shape=(n, MLP_OUTPUTS)).copy()
        r = np.ctypeslib.as_array(ctypes.cast(L.ga_records_reads(h), ctypes.POINTER(ctypes.c_int64)), shape=(n,)).copy()
        return f, y, r
    finally:
        L.ga_records_free(h)


def set_proposal(ctx, mlp, threshold=0.5, keep_top=1, max_clusters=0):
    """Prune seed clusters the network scores below threshold in later align calls
    on ctx (mlp None: the plain method); only each read's first max_clusters
    clusters are scored, later ones are removed (0: all scored)."""
    _check(N.lib().ga_context_set_proposal(ctx.h, mlp.h if mlp is not None else None, threshold, keep_top, max_clusters),
           "ga_context_set_proposal")
    ctx.proposal = mlp  # keeps the network alive while it is set
