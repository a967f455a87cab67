"""MI355X-native seed extension for GraphAligner-style long-read to graph alignment.

Python mirror of the reference boundary (src/GraphAlignerWrapper.h:52-60) over the
C ABI in include/graphaligner_amd.h; the work runs in the native library.
"""
from .api import (Graph, Context, Index, params_preset, extend_batch, extend_batch_noalloc, seeds, align_reads,  # noqa: F401
                  upload_reads, align_resident, order_seeds, align_one_way, AlignmentResults, Mlp,
                  mlp_num_params, cluster_records, set_proposal, MLP_FEATURES, MLP_OUTPUTS)
