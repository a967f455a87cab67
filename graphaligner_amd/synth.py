"""Synthetic workloads: seeded variation graphs (GFA) and reads (FASTA).

The generator is native (csrc/ga_synth.cpp); this module holds the workload
presets and a FASTA reader.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libga_synth.so")


class GraphConfig(ctypes.Structure):
    _fields_ = [("genomeLength", ctypes.c_int64), ("snpMin", ctypes.c_int64), ("snpMax", ctypes.c_int64),
                ("indelMin", ctypes.c_int64), ("indelMax", ctypes.c_int64), ("indelLenMin", ctypes.c_int64),
                ("indelLenMax", ctypes.c_int64), ("svEvery", ctypes.c_int64), ("svLenMin", ctypes.c_int64),
                ("svLenMax", ctypes.c_int64), ("nodeMax", ctypes.c_int32), ("seed", ctypes.c_uint64)]


class ReadConfig(ctypes.Structure):
    _fields_ = [("numReads", ctypes.c_int64), ("lengthModel", ctypes.c_int32), ("lenA", ctypes.c_int64),
                ("lenB", ctypes.c_int64), ("sigma", ctypes.c_double), ("lenMin", ctypes.c_int64),
                ("lenMax", ctypes.c_int64), ("errorMin", ctypes.c_double), ("errorMax", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


# Variation-graph workloads. "vg100k" is the 100 kbp / 1,000-read case with SNP
# and indel bubbles; "vg5m" the 5 Mbp graph with structural-variant bubbles and
# 20,000 log-normal reads at 12% error. The "mini" presets are small cases for
# parity tests.
PRESETS = {
    "vg100k": (dict(genomeLength=100_000, snpMin=100, snpMax=200, indelMin=500, indelMax=1500, indelLenMin=1,
                    indelLenMax=10, svEvery=0, svLenMin=50, svLenMax=5000, nodeMax=32, seed=1),
               dict(numReads=1000, lengthModel=0, lenA=5000, lenB=20000, sigma=0.0, lenMin=5000, lenMax=20000,
                    errorMin=0.10, errorMax=0.10, seed=1001)),
    "vg5m": (dict(genomeLength=5_000_000, snpMin=100, snpMax=200, indelMin=500, indelMax=1500, indelLenMin=1,
                  indelLenMax=10, svEvery=20_000, svLenMin=50, svLenMax=5000, nodeMax=32, seed=2),
             dict(numReads=20000, lengthModel=1, lenA=12000, lenB=0, sigma=0.6, lenMin=1000, lenMax=60000,
                  errorMin=0.12, errorMax=0.12, seed=1002)),
    # the batched mixed stream on the 5 Mbp graph: 100,000 reads, log-normal
    # lengths (median 10 kbp, sigma 0.9, clipped to 1-100 kbp), per-read error
    # drawn from 5-15 %
    "vg5m_stream": (dict(genomeLength=5_000_000, snpMin=100, snpMax=200, indelMin=500, indelMax=1500, indelLenMin=1,
                         indelLenMax=10, svEvery=20_000, svLenMin=50, svLenMax=5000, nodeMax=32, seed=2),
                    dict(numReads=100000, lengthModel=1, lenA=10000, lenB=0, sigma=0.9, lenMin=1000, lenMax=100000,
                         errorMin=0.05, errorMax=0.15, seed=1004)),
    "mini": (dict(genomeLength=20_000, snpMin=100, snpMax=200, indelMin=500, indelMax=1500, indelLenMin=1,
                  indelLenMax=10, svEvery=0, svLenMin=50, svLenMax=5000, nodeMax=32, seed=11),
             dict(numReads=40, lengthModel=0, lenA=500, lenB=3000, sigma=0.0, lenMin=500, lenMax=3000,
                  errorMin=0.10, errorMax=0.10, seed=1011)),
    "mini_sv": (dict(genomeLength=60_000, snpMin=100, snpMax=200, indelMin=500, indelMax=1500, indelLenMin=1,
                     indelLenMax=10, svEvery=6_000, svLenMin=50, svLenMax=1500, nodeMax=32, seed=12),
                dict(numReads=40, lengthModel=1, lenA=3000, lenB=0, sigma=0.6, lenMin=500, lenMax=8000,
                     errorMin=0.12, errorMax=0.12, seed=1012)),
}


class DbgConfig(ctypes.Structure):
    _fields_ = [("genomeLength", ctypes.c_int64), ("repeatCopies", ctypes.c_int64), ("familyLenMin", ctypes.c_int64),
                ("familyLenMax", ctypes.c_int64), ("copiesMin", ctypes.c_int64), ("copiesMax", ctypes.c_int64),
                ("divergenceMax", ctypes.c_double), ("tandemArrays", ctypes.c_int64), ("periodMin", ctypes.c_int64),
                ("periodMax", ctypes.c_int64), ("unitsMin", ctypes.c_int64), ("unitsMax", ctypes.c_int64),
                ("tandemDivergence", ctypes.c_double), ("k", ctypes.c_int32), ("seed", ctypes.c_uint64)]


# Cyclic de Bruijn graph workloads (aligned with preset -x dbg): a seeded genome
# with 300 inserted copies of repeat families (1-8 kbp, 2-20 copies each, each
# copy substituted at 0-3 %, either strand) and tandem arrays of period
# 10-200 bp, compacted into k = 63 unitigs with 62M overlaps. "dbg2m" is the
# 2 Mbp genome with 10,000 reads of 10-30 kbp at 8 % error; "mini_dbg" a small
# case for parity tests.
DBG_PRESETS = {
    "dbg2m": (dict(genomeLength=2_000_000, repeatCopies=300, familyLenMin=1000, familyLenMax=8000, copiesMin=2, copiesMax=20,
                   divergenceMax=0.03, tandemArrays=200, periodMin=10, periodMax=200, unitsMin=5, unitsMax=50,
                   tandemDivergence=0.01, k=63, seed=3),
              dict(numReads=10000, lengthModel=0, lenA=10000, lenB=30000, sigma=0.0, lenMin=10000, lenMax=30000,
                   errorMin=0.08, errorMax=0.08, seed=1003)),
    "mini_dbg": (dict(genomeLength=150_000, repeatCopies=40, familyLenMin=500, familyLenMax=3000, copiesMin=2, copiesMax=20,
                      divergenceMax=0.03, tandemArrays=20, periodMin=10, periodMax=200, unitsMin=5, unitsMax=50,
                      tandemDivergence=0.01, k=63, seed=13),
                 dict(numReads=60, lengthModel=0, lenA=1500, lenB=6000, sigma=0.0, lenMin=1500, lenMax=6000,
                      errorMin=0.08, errorMax=0.08, seed=1013)),
}


def _lib():
    if not os.path.exists(_LIB):
        raise RuntimeError("synthetic generator not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_LIB)
    L.ga_synth_variation_workload.restype = ctypes.c_int64
    L.ga_synth_variation_workload.argtypes = [ctypes.POINTER(GraphConfig), ctypes.POINTER(ReadConfig), ctypes.c_char_p]
    L.ga_synth_dbg_workload.restype = ctypes.c_int64
    L.ga_synth_dbg_workload.argtypes = [ctypes.POINTER(DbgConfig), ctypes.POINTER(ReadConfig), ctypes.c_char_p]
    return L


def generate(preset, prefix, overrides_graph=None, overrides_reads=None):
    """Write <prefix>.gfa and <prefix>.fa for a preset; returns the node count."""
    if preset in DBG_PRESETS:
        g, r = DBG_PRESETS[preset]
        g = dict(g, **(overrides_graph or {}))
        r = dict(r, **(overrides_reads or {}))
        n = _lib().ga_synth_dbg_workload(ctypes.byref(DbgConfig(**g)), ctypes.byref(ReadConfig(**r)), prefix.encode())
    else:
        g, r = PRESETS[preset]
        g = dict(g, **(overrides_graph or {}))
        r = dict(r, **(overrides_reads or {}))
        n = _lib().ga_synth_variation_workload(ctypes.byref(GraphConfig(**g)), ctypes.byref(ReadConfig(**r)), prefix.encode())
    if n < 0:
        raise RuntimeError("generator failed to write " + prefix)
    return n


def read_fasta(path):
    names, seqs = [], []
    with open(path) as f:
        name, parts = None, []
        for line in f:
            line = line.rstrip("\n")
            if line.startswith(">"):
                if name is not None:
                    names.append(name)
                    seqs.append("".join(parts))
                name, parts = line[1:].split()[0], []
            else:
                parts.append(line)
        if name is not None:
            names.append(name)
            seqs.append("".join(parts))
    return names, seqs
