"""exactAlignmentPart (GraphAligner::exactAlignmentPart, src/GraphAligner.h:403-457),
the seed loop's test whether an alignment passes through a seed, against the
reference's bisection restated literally (tests/emu/ga_pipeline_emu.cpp), on
random traces whose read positions rise by zero or one per item, as the trace
rebuild makes them: the same answer everywhere, the reference's assertion on a
trace whose last position does not exceed its first included."""
import ctypes

import pytest

import pipe_ffi as P


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_exact_alignment_part_equals_bisection(seed):
    L = P.lib()
    L.pipe_exact_part_check.restype = ctypes.c_int64
    L.pipe_exact_part_check.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    asserts, found = ctypes.c_int64(), ctypes.c_int64()
    bad = L.pipe_exact_part_check(seed, 400, ctypes.byref(asserts), ctypes.byref(found))
    assert bad == 0
    # both outcomes occur, and the single-item traces fire the first assertion
    assert found.value > 1000
    assert asserts.value > 0
