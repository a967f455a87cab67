"""ctypes binding of tests/_build/libga_pipe.so: the product's host pipeline
(index, seeding, seed loop, selection, GAF) over a backend that runs the gfx950
device code (extension core, seed lookup core) on host threads."""
import ctypes
import os

import numpy as np

_LIB = None
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class PipeOptions(ctypes.Structure):
    _fields_ = [("bandwidth", ctypes.c_int32), ("errorCost", ctypes.c_int32), ("xdrop", ctypes.c_int32),
                ("sloppy", ctypes.c_int32), ("maxCellsPerSlice", ctypes.c_int64),
                ("seedExtendDensity", ctypes.c_double), ("multimapScoreFraction", ctypes.c_double),
                ("seedDensity", ctypes.c_double), ("minClusterSize", ctypes.c_int64),
                ("deviceLoop", ctypes.c_int32), ("maxSpeculativeSeeds", ctypes.c_int32),
                ("minAlignmentScore", ctypes.c_double), ("deviceSeedOrder", ctypes.c_int32),
                ("cigarMatchMismatchMerge", ctypes.c_int32)]


# library variants: the default build, and one whose traceback recomputes every
# column of a node it enters (so the end-slice and boundary checks the default
# skips for the columns it does not read still run)
VARIANTS = {"": "libga_pipe.so", "fullcheck": "libga_pipe_fullcheck.so"}
_LIBS = {}


def lib(variant=""):
    if variant not in _LIBS:
        L = ctypes.CDLL(os.path.join(ROOT, "tests", "_build", VARIANTS[variant]))
        L.pipe_graph_load.restype = ctypes.c_void_p
        L.pipe_graph_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.pipe_graph_free.argtypes = [ctypes.c_void_p]
        L.pipe_index_build.restype = ctypes.c_void_p
        L.pipe_index_build.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.pipe_index_free.argtypes = [ctypes.c_void_p]
        L.pipe_index_maxcount.restype = ctypes.c_int64
        L.pipe_index_maxcount.argtypes = [ctypes.c_void_p]
        L.pipe_index_num_kmers.restype = ctypes.c_int64
        L.pipe_index_num_kmers.argtypes = [ctypes.c_void_p]
        L.pipe_get_seeds.restype = ctypes.c_int
        L.pipe_get_seeds.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.pipe_get_seeds_device_code.restype = ctypes.c_int
        L.pipe_get_seeds_device_code.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64, ctypes.c_double,
                                                 ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.pipe_sort_check.restype = ctypes.c_int
        L.pipe_sort_check.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
        L.pipe_align_reads.restype = ctypes.c_int
        L.pipe_align_reads.argtypes = [ctypes.c_void_p, ctypes.POINTER(PipeOptions), ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p),
                                       ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]
        L.pipe_free.argtypes = [ctypes.c_void_p]
        L.pipe_set_fault.argtypes = [ctypes.c_int]
        L.pipe_align_one_way.restype = ctypes.c_int
        L.pipe_align_one_way.argtypes = [ctypes.c_void_p, ctypes.POINTER(PipeOptions), ctypes.c_char_p, ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]
        L.pipe_set_linear_proposal.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_float, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_int]
        L.pipe_records.argtypes = [ctypes.c_int]
        L.pipe_records_take.restype = ctypes.c_int64
        L.pipe_records_take.argtypes = [ctypes.POINTER(ctypes.c_void_p)] * 3
        L.pipe_proposal_stats.argtypes = [ctypes.POINTER(ctypes.c_int64)]
        L.pipe_cluster_features.restype = ctypes.c_int64
        L.pipe_cluster_features.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                            ctypes.c_int64, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_float),
                                            ctypes.c_int64]
        _LIBS[variant] = L
    return _LIBS[variant]


class Graph:
    def __init__(self, path, variant=""):
        self.L = lib(variant)
        err = ctypes.create_string_buffer(512)
        self.h = self.L.pipe_graph_load(path.encode(), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.L.pipe_graph_free(self.h)
            self.h = None


class Index:
    def __init__(self, graph, k, w, discard):
        self.graph = graph
        self.L = graph.L
        self.h = self.L.pipe_index_build(graph.h, k, w, discard)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.pipe_index_free(self.h)
            self.h = None

    @property
    def max_count(self):
        return self.L.pipe_index_maxcount(self.h)

    @property
    def num_kmers(self):
        return self.L.pipe_index_num_kmers(self.h)

    def seeds_device_code(self, seq, density):
        """Ordered seeds through the device seeding code built for the host
        (lookupReadKmers + seedsOfRead); None on a reference assertion."""
        cap = max(1024, 64 * len(seq))
        out = np.zeros(10 * cap, dtype=np.int64)
        n = ctypes.c_int64()
        rc = self.L.pipe_get_seeds_device_code(self.h, seq.encode(), len(seq), density,
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), cap, ctypes.byref(n))
        if rc == 2:
            return None
        assert rc == 0, rc
        return out[: 10 * n.value].reshape(-1, 10)

    def seeds(self, seq, density, order=False):
        """Seeds as rows of 10 SeedHit fields; None when OrderSeeds fired a
        reference assertion (the read would be dropped)."""
        cap = max(1024, 64 * len(seq))
        out = np.zeros(10 * cap, dtype=np.int64)
        n = ctypes.c_int64()
        rc = self.L.pipe_get_seeds(self.h, seq.encode(), len(seq), density, int(order),
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), cap, ctypes.byref(n))
        if rc == 2:
            return None
        assert rc == 0
        return out[: 10 * n.value].reshape(-1, 10)


def options(oracle_params, density, device_loop=False, speculative=64, device_seed_order=False):
    p = oracle_params
    return PipeOptions(p.bandwidth, p.errorCost, p.xdrop, p.sloppy, p.maxCellsPerSlice, p.seedExtendDensity,
                       p.multimapScoreFraction, density, p.minClusterSize, int(device_loop), speculative,
                       p.minAlignmentScore, int(device_seed_order), int(p.cigarMatchMismatchMerge))


# execution modes of the seed loop: host-driven rounds one seed at a time, rounds
# with speculative extension of later seeds, the whole loop on the device
# (with devseeds: seed expansion and OrderSeeds through the backend's device code)
MODES = {"rounds": dict(device_loop=False, speculative=1), "rounds_spec": dict(device_loop=False, speculative=64),
         "device_loop": dict(device_loop=True, speculative=1), "batched": dict(device_loop=2, speculative=1),
         "rounds_devseeds": dict(device_loop=False, speculative=64, device_seed_order=True),
         "device_loop_devseeds": dict(device_loop=True, speculative=1, device_seed_order=True)}


def align_reads(index, oracle_params, density, names, seqs, device_loop=False, speculative=64, device_seed_order=False):
    n = len(seqs)
    na = (ctypes.c_char_p * n)(*[x.encode() for x in names])
    sa = (ctypes.c_char_p * n)(*[x.encode() for x in seqs])
    lens = np.array([len(s) for s in seqs], dtype=np.int64)
    out = ctypes.c_void_p()
    status = np.zeros(n, dtype=np.int32)
    stats = np.zeros(6, dtype=np.int64)
    opt = options(oracle_params, density, device_loop, speculative, device_seed_order)
    rc = index.L.pipe_align_reads(index.h, ctypes.byref(opt), n, na, sa, lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                  ctypes.byref(out), status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                  stats.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    if rc != 0:
        raise RuntimeError("pipeline failed (%d)" % rc)
    text = ctypes.string_at(out.value).decode()
    index.L.pipe_free(out)
    return text, status, dict(rounds=int(stats[0]), extensions=int(stats[1]), seeds=int(stats[2]), word_steps=int(stats[3]),
                              unused=int(stats[4]), launches=int(stats[5]))


def align_one_way(index, oracle_params, seq, seeds, select=False, mapq=False, speculative=64):
    """AlignOneWay with the given seeds (int64 rows of 10 SeedHit fields); the
    result as text in the layout of the oracle's align_one_way."""
    seeds = np.ascontiguousarray(seeds, dtype=np.int64).reshape(-1, 10)
    opt = options(oracle_params, 0.0, False, speculative)
    out = ctypes.c_void_p()
    rc = index.L.pipe_align_one_way(index.h, ctypes.byref(opt), seq.encode(), len(seq),
                                    seeds.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(seeds), int(select), int(mapq),
                                    ctypes.byref(out))
    if rc < 0:
        raise RuntimeError("pipeline failed")
    text = ctypes.string_at(out.value).decode()
    index.L.pipe_free(out)
    return text


def set_linear_proposal(weights, bias, threshold, keep_top=1, variant="", max_clusters=0):
    """Proposal with a logistic scorer over the 48 cluster features (None: off)."""
    L = lib(variant)
    if weights is None:
        L.pipe_set_linear_proposal(0, None, 0.0, 0.0, 1, 0)
        return
    w = np.ascontiguousarray(weights, dtype=np.float32)
    assert w.size == 48
    L.pipe_set_linear_proposal(1, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), bias, threshold, keep_top, max_clusters)


def records(on, variant=""):
    lib(variant).pipe_records(1 if on else 0)


def take_records(variant=""):
    """(features [n, 48], labels [n, 2], read index [n]) recorded since the last take."""
    L = lib(variant)
    f, y, r = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    n = L.pipe_records_take(ctypes.byref(f), ctypes.byref(y), ctypes.byref(r))
    assert n >= 0
    F = np.ctypeslib.as_array(ctypes.cast(f, ctypes.POINTER(ctypes.c_float)), shape=(max(n, 1) * 48,))[: n * 48].reshape(n, 48).copy()
    Y = np.ctypeslib.as_array(ctypes.cast(y, ctypes.POINTER(ctypes.c_float)), shape=(max(n, 1) * 2,))[: n * 2].reshape(n, 2).copy()
    R = np.ctypeslib.as_array(ctypes.cast(r, ctypes.POINTER(ctypes.c_int64)), shape=(max(n, 1),))[:n].copy()
    for p in (f, y, r):
        L.pipe_free(p)
    return F, Y, R


def proposal_stats(variant=""):
    out = np.zeros(3, dtype=np.int64)
    lib(variant).pipe_proposal_stats(out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return dict(clusters=int(out[0]), pruned_clusters=int(out[1]), pruned_seeds=int(out[2]))


def cluster_features(index, seq, seeds):
    """(cluster of every seed, features [clusters, 48]) of ordered seeds."""
    seeds = np.ascontiguousarray(seeds, dtype=np.int64).reshape(-1, 10)
    n = len(seeds)
    cl = np.zeros(max(n, 1), dtype=np.uint32)
    cap = max(n, 1)
    feats = np.zeros(cap * 48, dtype=np.float32)
    k = index.L.pipe_cluster_features(index.h, seq.encode(), len(seq), seeds.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), n,
                                      cl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                      feats.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), cap)
    assert k >= 0
    return cl[:n], feats[: k * 48].reshape(k, 48)
