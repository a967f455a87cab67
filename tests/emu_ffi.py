"""ctypes bindings of the host build of the device extension logic
(tests/_build/libga_emu.so, built from tests/emu/ga_emu.cpp). Test-only."""
import ctypes
import os

import numpy as np

from oracle_ffi import SliceSummary

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {"wave": os.path.join(_HERE, "_build", "libga_emu.so"), "thread": os.path.join(_HERE, "_build", "libga_thread.so")}
_loaded = {}


def lib(mode="wave"):
    if mode not in _loaded:
        L = ctypes.CDLL(_LIBS[mode])
        L.emu_graph_load.restype = ctypes.c_void_p
        L.emu_graph_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.emu_graph_free.argtypes = [ctypes.c_void_p]
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.emu_extend.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p,
                                 ctypes.c_int64, ctypes.c_int, ctypes.c_int64, i64p, ctypes.c_int64, i64p,
                                 ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(SliceSummary), ctypes.c_int64, i64p,
                                 ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32), ctypes.c_int64]
        u64 = ctypes.c_uint64
        L.emu_merge.argtypes = [u64, u64, ctypes.c_int32, u64, u64, ctypes.c_int32, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_int32)]
        L.emu_less_mask.restype = u64
        L.emu_less_mask.argtypes = [u64, u64, ctypes.c_int32, u64, u64, ctypes.c_int32]
        L.emu_max_x.restype = ctypes.c_int32
        L.emu_max_x.argtypes = [u64, u64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, u64]
        _loaded[mode] = L
    return _loaded[mode]


def _schedule_args(schedule):
    if schedule is None or len(schedule) == 0:
        return None, 0
    sch = np.ascontiguousarray(schedule, dtype=np.int32)
    _schedule_args.keep = sch
    return sch.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(sch)


class Graph:
    def __init__(self, path, mode="wave"):
        self.mode = mode
        err = ctypes.create_string_buffer(512)
        self.h = lib(mode).emu_graph_load(path.encode(), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib(self.mode).emu_graph_free(self.h)
            self.h = None


def extend(graph, params, seq, bigraph_node, offset, schedule=None):
    cap = 4 * len(seq) + 1024
    trace = np.zeros(4 * cap, dtype=np.int64)
    nt = ctypes.c_int64()
    score = ctypes.c_int32()
    scap = len(seq) // 64 + 4
    sums = (SliceSummary * scap)()
    ns = ctypes.c_int64()
    ws = ctypes.c_uint64()
    rc = lib(graph.mode).emu_extend(graph.h, params.bandwidth, params.maxCellsPerSlice, params.errorCost, params.xdrop, seq.encode(), len(seq),
                          bigraph_node, offset, trace.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), cap, ctypes.byref(nt),
                          ctypes.byref(score), sums, scap, ctypes.byref(ns), ctypes.byref(ws), *_schedule_args(schedule))
    rows = [(s.j, s.minScore, s.minScoreNode, s.minScoreNodeOffset, s.maxExactEndposScore, s.maxExactEndposNode,
             s.cellsProcessed, s.numNodes, s.itemsHash, s.scoresNotValid) for s in sums[: ns.value]]
    return rc, score.value, trace[: 4 * nt.value].reshape(-1, 4), rows
