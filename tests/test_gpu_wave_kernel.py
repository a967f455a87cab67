"""The wave-cooperative extension kernel: one extension per wavefront with the
64 lanes cooperating (ballots for searches, shuffles for merges and X-scores,
a node's columns one per lane), extensions pulled by a persistent work loop.
It is switched off: on MI355X it twice stopped producing output on the mini
workload's extensions (killed for silence, no fault reported). No cause has
been shown: on the CPU its lane code and its work loop, run on 64 threads per
wave that meet in every wave intrinsic, match the oracle on exactly those
extensions (tests/test_simt_lanes.py). Without a cause it is not compiled into
the product and not run on a GPU; the single-lane kernel runs instead."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.skip(reason="wave-cooperative extension kernel stopped producing output twice on MI355X (mini workload, "
                         "extensions of 513/280/2084/526 bases); no cause shown on the CPU, so it is not built or run")
def test_gpu_wave_cooperative_extension_kernel():
    """Would run the four mini extensions through the wave-cooperative kernel and
    compare status, score, trace and slice words with the oracle."""
    raise AssertionError("the wave-cooperative kernel is not built")
