"""ctypes bindings of the C ABI (include/graphaligner_amd.h).

The product path: every call goes into _build/libgraphaligner_amd.so. If the
library is missing the import fails loudly; GPU entry points fail with
GA_ENODEVICE when no HIP device is present.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libgraphaligner_amd.so")

GA_OK, GA_EXT_FAILED, GA_ASSERT, GA_EOVERFLOW, GA_EBADARG, GA_EIO, GA_ENODEVICE, GA_EHIP, GA_ENOMEM, GA_ESTALE, GA_ESCRATCH = range(11)
STATUS_NAMES = {0: "GA_OK", 1: "GA_EXT_FAILED", 2: "GA_ASSERT", 3: "GA_EOVERFLOW", 4: "GA_EBADARG", 5: "GA_EIO",
                6: "GA_ENODEVICE", 7: "GA_EHIP", 8: "GA_ENOMEM", 9: "GA_ESTALE", 10: "GA_ESCRATCH"}

# every entry point the header declares
EXPORTED = ["ga_params_preset", "ga_graph_load_gfa", "ga_graph_free", "ga_graph_num_nodes", "ga_graph_num_edges",
            "ga_graph_first_ambiguous", "ga_graph_export", "ga_context_create", "ga_context_free", "ga_extend_batch",
            "ga_extend_batch_scheduled",
            "ga_index_build", "ga_index_build_gpu", "ga_index_export", "ga_index_num_positions", "ga_index_free", "ga_index_max_count",
            "ga_index_num_kmers", "ga_seeds",
            "ga_align_reads", "ga_reads_upload", "ga_align_resident", "ga_context_synchronize", "ga_free",
            "ga_order_seeds", "ga_align_one_way", "ga_result_reads", "ga_result_status", "ga_result_seeds_extended",
            "ga_result_count", "ga_result_item", "ga_result_trace", "ga_select_alignments", "ga_add_mapping_qualities",
            "ga_gaf_line", "ga_result_free", "ga_extend_batch_noalloc", "ga_mlp_num_params", "ga_mlp_create",
            "ga_mlp_free", "ga_mlp_get_params", "ga_mlp_set_params", "ga_mlp_forward", "ga_mlp_gradients", "ga_mlp_train",
            "ga_cluster_records", "ga_records_count", "ga_records_features", "ga_records_labels", "ga_records_reads",
            "ga_records_free", "ga_context_set_proposal"]

READ_STATUS_NAMES = {0: "aligned", 1: "no seeds", 2: "assertion (read dropped)", 3: "no alignment", 4: "workspace overflow"}


class Params(ctypes.Structure):
    _fields_ = [("bandwidth", ctypes.c_int64), ("maxCellsPerSlice", ctypes.c_int64),
                ("preciseClippingIdentity", ctypes.c_double), ("xdrop", ctypes.c_int32), ("tryAllSeeds", ctypes.c_int32),
                ("seedExtendDensity", ctypes.c_double), ("seedClusterMinSize", ctypes.c_int64),
                ("multimapScoreFraction", ctypes.c_double), ("minAlignmentScore", ctypes.c_double),
                ("minimizerLength", ctypes.c_int32), ("minimizerWindowSize", ctypes.c_int32),
                ("minimizerSeedDensity", ctypes.c_double), ("minimizerDiscardMostNumerousFraction", ctypes.c_double),
                ("cigarMatchMismatchMerge", ctypes.c_int32), ("deviceSeedLoop", ctypes.c_int32),
                ("maxSpeculativeSeeds", ctypes.c_int32), ("deviceSeedOrder", ctypes.c_int32)]


class SliceSummary(ctypes.Structure):
    _fields_ = [("j", ctypes.c_int64), ("minScore", ctypes.c_int32), ("pad0", ctypes.c_int32),
                ("minScoreNode", ctypes.c_int64), ("minScoreNodeOffset", ctypes.c_int64),
                ("maxExactEndposScore", ctypes.c_int32), ("pad1", ctypes.c_int32), ("maxExactEndposNode", ctypes.c_int64),
                ("cellsProcessed", ctypes.c_int64), ("numNodes", ctypes.c_int64), ("itemsDigest", ctypes.c_uint64),
                ("scoresNotValid", ctypes.c_int32), ("pad2", ctypes.c_int32)]


class AlignStats(ctypes.Structure):
    _fields_ = [("lookupMs", ctypes.c_double), ("extendKernelMs", ctypes.c_double), ("hostMs", ctypes.c_double),
                ("totalMs", ctypes.c_double), ("rounds", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("extensions", ctypes.c_int64), ("wordSteps", ctypes.c_int64), ("seeds", ctypes.c_int64),
                ("reruns", ctypes.c_int64), ("lookupCallMs", ctypes.c_double), ("extendCallMs", ctypes.c_double),
                ("unusedExtensions", ctypes.c_int64), ("seedingMs", ctypes.c_double), ("pickMs", ctypes.c_double),
                ("replayMs", ctypes.c_double), ("gafMs", ctypes.c_double), ("gafBytes", ctypes.c_int64),
                ("persistentWorkers", ctypes.c_int64), ("tuneMs", ctypes.c_double),
                ("proposalClusters", ctypes.c_int64), ("prunedClusters", ctypes.c_int64), ("prunedSeeds", ctypes.c_int64),
                ("proposalMs", ctypes.c_double)]


class MlpTrainParams(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_int32), ("batch", ctypes.c_int32), ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("seed", ctypes.c_uint64)]


class SeedHit(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int64) for f in ("nodeID", "nodeOffset", "seqPos", "matchLen", "reverse", "alignmentGraphNodeId",
                                              "alignmentGraphNodeOffset", "rawSeedGoodness", "seedGoodness", "seedClusterSize")]


class AlignmentItem(ctypes.Structure):
    _fields_ = [("alignmentStart", ctypes.c_int64), ("alignmentEnd", ctypes.c_int64), ("alignmentScore", ctypes.c_int64),
                ("alignmentXScore", ctypes.c_double), ("seedGoodness", ctypes.c_int64), ("mappingQuality", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("traceLength", ctypes.c_int64)]


class TraceItem(ctypes.Structure):
    _fields_ = [("node", ctypes.c_int64), ("nodeOffset", ctypes.c_int64), ("seqPos", ctypes.c_int64),
                ("nodeSwitch", ctypes.c_int8), ("sequenceCharacter", ctypes.c_char), ("graphCharacter", ctypes.c_char),
                ("pad", ctypes.c_char * 5)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("graphaligner_amd native library missing (" + LIB_PATH + "): run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        L.ga_params_preset.argtypes = [ctypes.c_char_p, ctypes.POINTER(Params)]
        L.ga_graph_load_gfa.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
        L.ga_graph_free.argtypes = [vp]
        for f in ("ga_graph_num_nodes", "ga_graph_num_edges", "ga_graph_first_ambiguous"):
            getattr(L, f).restype = i64
            getattr(L, f).argtypes = [vp]
        L.ga_graph_export.argtypes = [vp] + [vp] * 13
        L.ga_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
        L.ga_context_free.argtypes = [vp]
        L.ga_context_synchronize.argtypes = [vp]
        L.ga_extend_batch.argtypes = [vp, vp, ctypes.POINTER(Params), i64, ctypes.POINTER(ctypes.c_char_p), vp, vp, vp, i64, vp,
                                      vp, vp, vp, ctypes.POINTER(SliceSummary), i64, vp, ctypes.POINTER(ctypes.c_double)]
        L.ga_extend_batch_scheduled.argtypes = [vp, vp, ctypes.POINTER(Params), i64, ctypes.POINTER(ctypes.c_char_p), vp, vp, vp,
                                                vp, vp, vp, i64, vp, vp, vp, vp, ctypes.POINTER(SliceSummary), i64, vp,
                                                ctypes.POINTER(ctypes.c_double)]
        L.ga_index_build.argtypes = [vp, ctypes.POINTER(Params), ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
        L.ga_index_free.argtypes = [vp]
        L.ga_index_build_gpu.argtypes = [vp, vp, ctypes.POINTER(Params), ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
        L.ga_index_export.argtypes = [vp, vp, vp, vp, vp]
        L.ga_index_num_positions.restype = i64
        L.ga_index_num_positions.argtypes = [vp]
        L.ga_index_max_count.restype = i64
        L.ga_index_max_count.argtypes = [vp]
        L.ga_index_num_kmers.restype = i64
        L.ga_index_num_kmers.argtypes = [vp]
        L.ga_seeds.argtypes = [vp, vp, vp, ctypes.c_double, i32, i64, ctypes.POINTER(ctypes.c_char_p), vp,
                               ctypes.POINTER(vp), vp]
        L.ga_align_reads.argtypes = [vp, vp, vp, ctypes.POINTER(Params), i64, ctypes.POINTER(ctypes.c_char_p),
                                     ctypes.POINTER(ctypes.c_char_p), vp, ctypes.POINTER(vp), vp, ctypes.POINTER(AlignStats)]
        L.ga_reads_upload.argtypes = [vp, i64, ctypes.POINTER(ctypes.c_char_p), vp]
        L.ga_align_resident.argtypes = [vp, vp, vp, ctypes.POINTER(Params), i64, i64, ctypes.POINTER(ctypes.c_char_p),
                                        ctypes.POINTER(vp), vp, ctypes.POINTER(AlignStats)]
        L.ga_free.argtypes = [vp]
        L.ga_order_seeds.argtypes = [vp, ctypes.POINTER(SeedHit), i64]
        L.ga_align_one_way.argtypes = [vp, vp, ctypes.POINTER(Params), i64, ctypes.POINTER(ctypes.c_char_p), vp,
                                       ctypes.POINTER(SeedHit), vp, ctypes.POINTER(vp)]
        L.ga_result_reads.restype = i64
        L.ga_result_reads.argtypes = [vp]
        L.ga_result_status.restype = i32
        L.ga_result_status.argtypes = [vp, i64]
        L.ga_result_seeds_extended.restype = i64
        L.ga_result_seeds_extended.argtypes = [vp, i64]
        L.ga_result_count.restype = i64
        L.ga_result_count.argtypes = [vp, i64]
        L.ga_result_item.argtypes = [vp, i64, i64, ctypes.POINTER(AlignmentItem)]
        L.ga_result_trace.argtypes = [vp, i64, i64, ctypes.POINTER(TraceItem), i64]
        L.ga_select_alignments.argtypes = [vp, i64, ctypes.POINTER(Params), i32]
        L.ga_add_mapping_qualities.argtypes = [vp, i64]
        L.ga_gaf_line.argtypes = [vp, vp, i64, i64, ctypes.c_char_p, i64, i32, ctypes.c_char_p, i64, ctypes.POINTER(i64)]
        L.ga_result_free.argtypes = [vp]
        L.ga_extend_batch_noalloc.argtypes = [vp, vp, ctypes.POINTER(Params), i64, ctypes.POINTER(ctypes.c_char_p), vp, vp, vp,
                                              i64, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
        L.ga_mlp_num_params.restype = i64
        L.ga_mlp_num_params.argtypes = []
        L.ga_mlp_create.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.ga_mlp_free.argtypes = [vp]
        L.ga_mlp_get_params.argtypes = [vp, vp]
        L.ga_mlp_set_params.argtypes = [vp, vp]
        L.ga_mlp_forward.argtypes = [vp, i64, vp, vp]
        L.ga_mlp_gradients.argtypes = [vp, i64, vp, vp, ctypes.POINTER(ctypes.c_double), vp]
        L.ga_mlp_train.argtypes = [vp, i64, vp, vp, ctypes.POINTER(MlpTrainParams), vp, ctypes.POINTER(ctypes.c_double)]
        L.ga_cluster_records.argtypes = [vp, vp, vp, ctypes.POINTER(Params), i64, i64, ctypes.POINTER(vp)]
        L.ga_records_count.restype = i64
        L.ga_records_count.argtypes = [vp]
        for f in ("ga_records_features", "ga_records_labels", "ga_records_reads"):
            getattr(L, f).restype = vp
            getattr(L, f).argtypes = [vp]
        L.ga_records_free.argtypes = [vp]
        L.ga_context_set_proposal.argtypes = [vp, vp, ctypes.c_double, i32, i32]
        _lib = L
    return _lib
